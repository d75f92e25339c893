"""GPU adjoint (path replay) of the render against finite differences of the
oracle: the oracle re-evaluates the estimator in double precision with the
base render's discrete decisions frozen (oracle/diff.cpp), central differences
with h = 1e-4; gradients of the L2 image loss against a target rendered with
parameters perturbed by up to +-20 % (seeded).  Bound: 2 % relative error per
parameter (plus an absolute floor of 1e-3 of the family's largest gradient)."""
import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import lib, scenes

pytestmark = pytest.mark.gpu

H_FD = 1e-4
REL_TOL = 0.02


def _setup(ctx, width=128, height=128, spp=2, freq=40):
    s = scenes.c4_scene(width=width, height=height, spp=spp, freq=freq)
    r = lib.DiffRenderer(ctx, s)
    alpha, etak, env = r.params()
    g = np.random.default_rng(2026)
    ta = alpha * (1 + 0.2 * g.uniform(-1, 1, alpha.shape)).astype(np.float32)
    tek = etak * (1 + 0.2 * g.uniform(-1, 1, etak.shape)).astype(np.float32)
    tenv = env * (1 + 0.2 * g.uniform(-1, 1, env.shape)).astype(np.float32)
    r.set_params(ta, tek, tenv)
    r.forward()
    target = r.film_rgb().astype(np.float32)
    r.set_params(alpha, etak, env)
    return s, r, (alpha, etak, env), target


def test_adjoint_matches_finite_differences(ctx):
    s, r, (alpha, etak, env), target = _setup(ctx)
    print("scene", s.n_shapes, "shapes", flush=True)
    r.forward()
    r.set_target(target)
    loss = r.loss()
    r.backward()
    ga, ge, genv = r.gradients()
    D = po.DiffOracle(s)
    base = [alpha.astype(np.float64), etak.astype(np.float64), env.astype(np.float64)]
    loss_o = D.loss(*base, target)
    print("loss gpu %.6g oracle %.6g" % (loss, loss_o), flush=True)
    assert loss == pytest.approx(loss_o, rel=1e-3)
    types = np.asarray(s.materials, np.float32).reshape(-1, 9)[:, 0].view(np.uint32)
    # only conductors read eta/k as RGB
    assert np.all(ge[types != scenes.MAT_CONDUCTOR] == 0)
    assert np.any(ge[types == scenes.MAT_CONDUCTOR] != 0)
    g = np.random.default_rng(5)
    picks = []
    for fam, grad, n in (("alpha", ga, 12), ("eta_k", ge.ravel(), 10), ("env", genv.ravel(), 10)):
        live = np.flatnonzero(np.abs(grad) > 1e-6 * np.abs(grad).max())
        for k in g.choice(live, size=min(n, len(live)), replace=False):
            picks.append((fam, int(k), float(grad[k]), float(np.abs(grad).max())))
    assert len(picks) >= 24
    bad = []
    for fam, k, adj, gmax in picks:
        fd = []
        for sign in (1, -1):
            p = [b.copy() for b in base]
            arr = {"alpha": p[0], "eta_k": p[1].reshape(-1), "env": p[2].reshape(-1)}[fam]
            arr[k] += sign * H_FD
            fd.append(D.loss(*p, target))
        fd = (fd[0] - fd[1]) / (2 * H_FD)
        err = abs(adj - fd)
        ok = err <= REL_TOL * abs(fd) + 1e-3 * gmax
        print("%-6s %6d adjoint %+.6e fd %+.6e rel %.2e %s" % (fam, k, adj, fd, err / max(abs(fd), 1e-30),
                                                             "" if ok else "<-- out of bound"), flush=True)
        if not ok:
            bad.append((fam, k, adj, fd))
    assert not bad, bad
