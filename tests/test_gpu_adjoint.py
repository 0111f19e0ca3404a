"""S(k)-misfit loss and its adjoint on the device (differentiation form):
N = 32,768 atoms of a jittered fcc lattice (21^3 cells, seeded subset),
L = 30.3, 2,000 vectors in 50 shells with k <= 10 (bins 47..96, the first
40 vectors of each reference shell), target S(k) from a second seeded
configuration.  Checked against the FP64 oracle gradient and against
central finite differences (h = 1e-4) of the FP64 oracle loss."""
import os

import numpy as np
import pytest

from amdat_amd import _abi, synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

orc.set_threads(min(16, os.cpu_count() or 1))


def shells_c4(qvec):
    off, n = qvec
    rows = [n[off[b - 1]:off[b - 1] + 40] for b in range(47, 97)]
    so = np.arange(0, 2001, 40, dtype=np.int32)
    return so, np.ascontiguousarray(np.concatenate(rows))


@pytest.fixture(scope="module")
def c4(qvec):
    print("case: adjoint N=32768 L=30.3 K=2000", flush=True)
    # the device takes the length scale as FP32: give the oracle the same value
    L = float(np.float32(30.3))
    a = synth.fcc_lattice(21, L, jitter=0.1, n_keep=32768, seed=synth.MASTER_SEED + 3)
    b = synth.fcc_lattice(21, L, jitter=0.2, n_keep=32768, seed=synth.MASTER_SEED + 103)
    pos = np.stack([a.x[0], a.y[0], a.z[0]], 1).astype(np.float32)
    posb = np.stack([b.x[0], b.y[0], b.z[0]], 1).astype(np.float64)
    off, n = shells_c4(qvec)
    w = np.ones(50)
    _, target, _ = orc.sk_loss_grad_f64(posb, off, n, L, w, np.zeros(50), grad=False)
    loss64, S64, g64 = orc.sk_loss_grad_f64(pos.astype(np.float64), off, n, L, w, target)
    return pos, off, n, L, w, target, loss64, S64, g64


def test_adjoint_matches_fp64_and_fd(ctx, c4):
    import torch
    from amdat_amd import torch_sk
    pos, off, n, L, w, target, loss64, S64, g64 = c4
    sh = _abi.Shells(ctx, off, n)
    p = torch.tensor(pos, device="cuda:0", requires_grad=True)
    loss = torch_sk.structure_factor_loss(p, sh, 0, 49, L, w, target)
    loss.backward()
    g = p.grad.double().cpu().numpy()
    S = torch_sk.StructureFactorLoss  # noqa: F841
    rel_loss = abs(loss.item() - loss64) / abs(loss64)
    gscale = np.abs(g64).max()
    rel_g = np.abs(g - g64).max() / gscale
    print("adjoint: loss rel err %.3e, grad max err / max|g| %.3e" % (rel_loss, rel_g))
    assert rel_loss < 1e-6
    assert rel_g < 1e-5
    # central finite differences of the FP64 loss on a seeded sample of coordinates
    rng = np.random.default_rng(7)
    h = 1e-4
    base = pos.astype(np.float64)
    for j, c in zip(rng.choice(len(pos), 4, replace=False), rng.integers(0, 3, 4)):
        p1 = base.copy(); p1[j, c] += h
        p0 = base.copy(); p0[j, c] -= h
        fd = (orc.sk_loss_grad_f64(p1, off, n, L, w, target, grad=False)[0]
              - orc.sk_loss_grad_f64(p0, off, n, L, w, target, grad=False)[0]) / (2 * h)
        print("  coord (%d,%d): adjoint %.8e fd %.8e" % (j, c, g[j, c], fd))
        assert abs(g[j, c] - fd) <= 1e-5 * max(abs(fd), 1e-2 * gscale)


def test_loss_only_and_errors(ctx, c4):
    import torch
    pos, off, n, L, w, target, loss64, S64, g64 = c4
    sh = _abi.Shells(ctx, off, n)
    t = torch.tensor(pos, device="cuda:0")
    cols = [t[:, c].contiguous() for c in range(3)]
    loss, S = _abi.sk_loss_grad(ctx, len(pos), *[c.data_ptr() for c in cols], sh, 0, 49, L, w, target)
    assert np.max(np.abs(S - S64) / np.abs(S64)) < 1e-6
    with pytest.raises(_abi.AmdatError) as e:
        _abi.sk_loss_grad(ctx, len(pos), *[c.data_ptr() for c in cols], sh, 0, 50, L, np.ones(51), np.ones(51))
    assert e.value.code == -4
