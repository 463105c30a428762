"""Adjoint of the p2d step (galhip_adj_*) and the gradients built on it.

No counterpart in the reference; the checks are those that pin a
derivative without one:
  - the transposed field solve against the forward solve: <K rho, g> =
    <rho, K^T g> for random rho and g (FP64, 1e-12 relative);
  - the whole backward pass against central finite differences of the
    FP64 forward run: directional derivatives in the initial velocities,
    and dL/dQ, dL/dmd (the disc's Toomre Q and mass);
  - the FP32 forward's gradient against the FP64 one;
  - the torch.autograd.Function harness against the same numbers.
"""
import numpy as np
import pytest

from galaxy_amd import adjoint, configs
from galaxy_amd.lib15 import P2DGrid

pytestmark = pytest.mark.gpu

W = configs.Workload("p2d-64x128-adjoint-test", nr=64, na=128, hash=8, n=20000, steps=12,
                     dt=0.05, Q=1.2, seed=5)


def _say(*a):
    print("[case]", *a, flush=True)


def test_field_transpose():
    """<K rho, g> = <rho, K^T g>: K = galhip_findf on a mass grid set with
    galhip_set_mass, K^T = galhip_adj_field, both with the same ranges."""
    _say("adj_field transpose")
    kw = configs.grid_kwargs(W, W.n)
    g = P2DGrid(precision="fp64", **kw)
    g.greenm()
    rng = np.random.default_rng(3)
    mfx = rng.integers(0, 1 << 40, size=(1, g.mesh), dtype=np.int64)
    g.set_mass(mfx=mfx, jrad=np.array([1], np.int32), krad=np.array([g.nr], np.int32))
    g.findf(False)
    rho = g.get_rho()
    F = g.get_field()
    gg = rng.standard_normal((2, g.mesh))
    rb = g.adj_field(gg)
    lhs = float(np.dot(F[0], gg[0]) + np.dot(F[1], gg[1]))
    rhs = float(np.dot(rho, rb))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1e-300), (lhs, rhs)
    # with a filtered harmonic and a partial source range too
    g.close()
    g = P2DGrid(precision="fp64", skip=(3,), **kw)
    g.greenm()
    mfx[0, : 5 * g.na] = 0
    g.set_mass(mfx=mfx, jrad=np.array([6], np.int32), krad=np.array([g.nr - 3], np.int32))
    g.findf(False)
    rho, F = g.get_rho(), g.get_field()
    rb = g.adj_field(gg)
    lhs = float(np.dot(F[0], gg[0]) + np.dot(F[1], gg[1]))
    jrs, krs = g.ranges()[:2]
    sel = np.zeros(g.mesh, bool)
    sel[(jrs - 1) * g.na: krs * g.na] = True
    rhs = float(np.dot(rho[sel], rb[sel]))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    g.close()


def _fd(f, h):
    return (f(h) - f(-h)) / (2 * h)


def test_gradient_matches_finite_differences():
    """FP64 forward; 12 steps with checkpoints every 5 (a partial last
    segment).  Directional derivative of L in the initial velocities along
    two random directions, and dL/dQ, dL/dmd, against central differences."""
    _say("adjoint vs finite differences", W.n, W.steps)
    m = adjoint.DiscModel(W, precision="fp64")
    Q, md = W.Q, W.md
    gr = m.gradient(Q, md, W.steps, every=5)
    vx, vy = m.velocities(Q, md)
    rng = np.random.default_rng(11)
    scale = float(np.sqrt(np.mean(vx * vx + vy * vy)))
    for k in range(2):
        dx, dy = rng.standard_normal(m.n), rng.standard_normal(m.n)
        ad = float(np.dot(gr["dvx"], dx) + np.dot(gr["dvy"], dy))
        h = 1e-6 * scale
        fd = _fd(lambda e: m.loss_v(vx + e * dx, vy + e * dy, md, W.steps), h)
        _say("direction %d: adjoint %.9e finite difference %.9e" % (k, ad, fd))
        assert abs(ad - fd) <= 1e-4 * abs(fd), (ad, fd)
    fq = _fd(lambda e: m.loss(Q + e, md, W.steps), 1e-6)
    _say("dL/dQ: adjoint %.9e finite difference %.9e" % (gr["dQ"], fq))
    assert abs(gr["dQ"] - fq) <= 1e-4 * abs(fq)
    fm = _fd(lambda e: m.loss(Q, md + e, W.steps), 1e-6)
    _say("dL/dmd: adjoint %.9e finite difference %.9e" % (gr["dmd"], fm))
    assert abs(gr["dmd"] - fm) <= 1e-4 * abs(fm)


def test_fp32_gradient_near_fp64():
    _say("fp32 vs fp64 gradient")
    g64 = adjoint.DiscModel(W, precision="fp64").gradient(W.Q, W.md, W.steps, every=5)
    g32 = adjoint.DiscModel(W, precision="fp32").gradient(W.Q, W.md, W.steps, every=5)
    v64 = np.concatenate([g64["dvx"], g64["dvy"]])
    v32 = np.concatenate([g32["dvx"], g32["dvy"]])
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    _say("relative difference of the velocity gradients %.3e, dQ %.6e vs %.6e"
         % (err, g32["dQ"], g64["dQ"]))
    assert err < 1e-3
    assert abs(g32["dQ"] - g64["dQ"]) <= 1e-3 * abs(g64["dQ"])


def test_torch_harness():
    """torch autograd through the Function gives the model's dQ and dmd."""
    import torch
    _say("torch harness")
    m = adjoint.DiscModel(W, precision="fp64")
    Q = torch.tensor(W.Q, dtype=torch.float64, requires_grad=True)
    md = torch.tensor(W.md, dtype=torch.float64, requires_grad=True)
    L = adjoint.torch_loss(m, Q, md, W.steps, every=5)
    L.backward()
    gr = m.gradient(W.Q, W.md, W.steps, every=5)
    assert abs(L.item() - gr["loss"]) <= 1e-12 * gr["loss"]
    assert abs(Q.grad.item() - gr["dQ"]) <= 1e-8 * abs(gr["dQ"]), (Q.grad.item(), gr["dQ"])
    assert abs(md.grad.item() - gr["dmd"]) <= 1e-8 * abs(gr["dmd"]), (md.grad.item(), gr["dmd"])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_device_seed_matches_host_loss(prec):
    """galhip_adj_seed_a2 against a2_loss on the unloaded particles: L and
    A_2 to 1e-12, the seeded adjoints (read back in load order) to 1e-12 of
    their maximum.  Then the backward pass from the device seed with the
    parameter sums kept on the device, against the pass from adj_set of the
    host seed with the sums taken step by step: initial adjoints and gpar
    to 1e-10."""
    _say("device A_2 seed", prec)
    m = adjoint.DiscModel(W, precision=prec)
    vx, vy = m.velocities(W.Q, W.md)
    res = []
    for dev in (True, False):
        g = m.grid(W.md)
        tr = adjoint.Trajectory(g, W.steps, 5)
        out = tr.forward(m.x0, m.y0, vx, vy, m.iz0)
        L, gx, gy = adjoint.a2_loss(out["x"], out["y"], m.ls)
        if dev:
            Ld, a2 = g.adj_seed_a2(m.ls, 1.0, 3.0)
            assert abs(Ld - L) <= 1e-12 * L, (Ld, L)
            assert abs(abs(a2) ** 2 - L) <= 1e-12 * L
            lam = g.adj_get()
            for k, ref in ((0, gx), (1, gy)):
                err = np.abs(lam[k] - ref).max()
                assert err <= 1e-12 * np.abs(ref).max(), (k, err)
            assert not lam[2:].any()
            lam0, gpar = tr.backward()
        else:
            lam = np.zeros((4, m.n))
            lam[0], lam[1] = gx, gy
            g.adj_set(lam)
            gpar = np.zeros(2)
            base = tr.nck
            for k in range(tr.nck - 1, -1, -1):
                mm = min(tr.every, tr.nsteps - k * tr.every)
                g.adj_restore(k)
                for j in range(mm):
                    g.adj_record(base + j)
                    g.run(1)
                for j in range(mm - 1, -1, -1):
                    g.adj_restore(base + j)
                    g.findf(False)
                    g.adj_step(gpar)
            assert np.array_equal(g.adj_get_par(), gpar)
            lam0 = g.adj_get()
        g.close()
        res.append((lam0, gpar))
    (a, pa), (b, pb) = res
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    assert np.all(np.abs(pa - pb) <= 1e-10 * np.abs(pb)), (pa, pb)


def test_model_grid_reuse():
    """DiscModel keeps its grid between gradient calls (galhip_set_pmass and
    a new halo table when md changes): a second call at another md against
    a fresh model's call.  The forward run is deterministic, so the loss is
    bit-identical; the adjoint scatter sums with FP64 atomics in no fixed
    order, so the gradients agree to rounding (1e-10 here)."""
    _say("model grid reuse")
    md2 = 1.15 * W.md
    m = adjoint.DiscModel(W, precision="fp32")
    m.gradient(W.Q, W.md, W.steps, every=5)
    a = m.gradient(1.1 * W.Q, md2, W.steps, every=5)
    b = adjoint.DiscModel(W, precision="fp32").gradient(1.1 * W.Q, md2, W.steps, every=5)
    assert a["loss"] == b["loss"], (a["loss"], b["loss"])
    for k in ("dQ", "dmd"):
        assert abs(a[k] - b[k]) <= 1e-10 * abs(b[k]), (k, a[k], b[k])
    for k in ("dvx", "dvy", "gpar"):
        d = np.abs(a[k] - b[k]).max()
        _say("%s: max |reused - fresh| %.3g of max %.3g" % (k, d, np.abs(b[k]).max()))
        assert d <= 1e-10 * np.abs(b[k]).max(), k
    m.close()
