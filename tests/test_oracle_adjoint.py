"""Pins of the oracle's S(k)-misfit loss and gradient: central finite
differences of the loss (FP64, h = 1e-4), and the closed form of S."""
import numpy as np

from amdat_amd import synth
from oracle import oracle as orc
from tests.conftest import ref_shells


def test_gradient_matches_central_differences(qvec):
    off, n = ref_shells(qvec, 10, 24)
    L = 6.0
    t = synth.fcc_lattice(3, L, jitter=0.1, seed=5)
    pos = np.stack([t.x[0], t.y[0], t.z[0]], 1).astype(np.float64)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, len(off) - 1)
    target = rng.uniform(0.5, 1.5, len(off) - 1)
    loss, S, g = orc.sk_loss_grad_f64(pos, off, n, L, w, target)
    h = 1e-4
    for j, c in [(0, 0), (5, 1), (17, 2), (40, 0), (77, 2), (100, 1)]:
        p1 = pos.copy(); p1[j, c] += h
        p0 = pos.copy(); p0[j, c] -= h
        fd = (orc.sk_loss_grad_f64(p1, off, n, L, w, target, grad=False)[0]
              - orc.sk_loss_grad_f64(p0, off, n, L, w, target, grad=False)[0]) / (2 * h)
        assert abs(fd - g[j, c]) <= 1e-6 * max(abs(g[j, c]), 1e-3 * np.abs(g).max()), (j, c, fd, g[j, c])
    # S matches the structure-factor oracle on the same configuration
    rho = orc.rho_f64(*pos.T.astype(np.float32), n, L)
    S2 = orc.sk_f64(rho[None], rho[None], np.array([len(pos)]), off)
    assert np.allclose(S, S2, rtol=1e-6)
    # translation invariance: the gradient sums to zero
    assert np.allclose(g.sum(0), 0.0, atol=1e-9 * np.abs(g).max() * len(pos))


def test_zero_misfit_zero_gradient(qvec):
    off, n = ref_shells(qvec, 4, 12)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 5.0, size=(30, 3))
    w = np.ones(len(off) - 1)
    _, S, _ = orc.sk_loss_grad_f64(pos, off, n, 5.0, w, np.zeros(len(off) - 1), grad=False)
    loss, _, g = orc.sk_loss_grad_f64(pos, off, n, 5.0, w, S)
    assert loss == 0.0 and np.all(g == 0.0)
