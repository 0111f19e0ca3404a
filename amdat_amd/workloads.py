"""The benchmark workloads, shared by bench.py and the full-size GPU tests
so that both run on the same inputs.

precision(): the 64,000-atom bead-spring melt (400 chains x 160 beads,
    L = 42.2, `exponential 10 100 1.2 0 0 1`, T = 1,001, sigma = 0.02),
    F/S shells bins 2,4,...,200 of the box-L grid (100 shells, 28,181
    vectors), Fs shells bins 2..100 at max_length_scale 21.05 (99 shells,
    26,243 vectors), the fullblock = 0 pairs (1,046).
batch_members(): the batched form, 256 trajectories with N uniform in
    [500, 20,000], density U[1.0, 1.2], `exponential 4 50 1.25` (T = 201),
    Gaussian steps sigma U[0.01, 0.1], generated on the device.
"""
import time

import numpy as np

from . import _abi, synth

PRECISION = dict(n_chains=400, beads=160, L=42.2, nexp=10, nes=100, base=1.2, sigma=0.02, ell_fs=21.05)
PRECISION_SMALL = dict(n_chains=40, beads=160, L=19.6, nexp=4, nes=25, base=1.2, sigma=0.02, ell_fs=9.8)

BATCH = dict(n_traj=256, n_min=500, n_max=20000, nexp=4, nes=50, sk_bins=(2, 61), fs_k=7.25)


def precision(small=False, log=None):
    """Returns cfg, trajectory (host), F/S shells (off, n), Fs shells
    (off, n), pairs (tg, t0, t1, n_timegaps)."""
    cfg = dict(PRECISION_SMALL if small else PRECISION)
    T = cfg["nes"] * cfg["nexp"] + 1
    t0 = time.time()
    tr = synth.chain_melt(cfg["n_chains"], cfg["beads"], cfg["L"], T, cfg["sigma"], seed=synth.MASTER_SEED + 1)
    if log:
        log("generated trajectory N=%d T=%d in %.1fs" % (tr.N, tr.T, time.time() - t0))
    # F/S shells: bins 2,4,...,200 of the box-L grid; Fs shells: bins 2..100.
    offA, nA = _abi.shells_generate(2, 199, 300, seed=synth.MASTER_SEED + 11)
    keep = list(range(0, 199, 2))
    rows = [nA[offA[s]:offA[s + 1]] for s in keep]
    offF = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    nF = np.concatenate(rows).astype(np.int32)
    offS, nS = _abi.shells_generate(2, 99, 300, seed=synth.MASTER_SEED + 12)
    tg, t0_, t1_ = _abi.ts_pairs(cfg["nexp"], cfg["nes"], 0, 0, 0)
    ntg = _abi.ts_n_timegaps(cfg["nexp"], cfg["nes"], 0)
    return cfg, tr, (offF, nF), (offS, nS), (tg, t0_, t1_, ntg)


def batch_members(ctx, n_traj, n_min, n_max, nexp, nes, seed):
    """Batched workload on the device: N_i ~ U{n_min..n_max}, density
    U[1.0, 1.2], L_i = (N_i/rho_i)^(1/3), uniform frame 0, Gaussian steps
    sigma_i ~ U[0.01, 0.1]; FP64 unwrapped positions stored as FP32 wrapped
    coordinates + int16 images (frame-major SoA), wrapped without copies.
    Returns (trajectory handles, the device columns that back them, box
    lengths)."""
    import torch
    rng = np.random.default_rng(seed)
    T = nes * nexp + 1
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    members, keep, Ls = [], [], []
    for i in range(n_traj):
        N = int(rng.integers(n_min, n_max + 1))
        L = float(np.float32((N / rng.uniform(1.0, 1.2)) ** (1.0 / 3.0)))
        sigma = float(rng.uniform(0.01, 0.1))
        u = torch.empty((T, N, 3), dtype=torch.float64, device="cuda:0")
        u[0].uniform_(0.0, L, generator=gen)
        u[1:].normal_(0.0, sigma, generator=gen)
        u = torch.cumsum(u, dim=0)
        img = torch.floor(u / L)
        w = (u - L * img).to(torch.float32)
        hit = w >= np.float32(L)
        w[hit] = 0.0
        img[hit] += 1.0
        cols = [w[..., c].contiguous() for c in range(3)] + [img[..., c].to(torch.int16).contiguous() for c in range(3)]
        del u, img, w, hit
        box = np.full((T, 3), L, np.float32)
        members.append(_abi.Trajectory.from_device(ctx, N, T, [c.data_ptr() for c in cols], box))
        keep.append(cols)
        Ls.append(L)
    torch.cuda.synchronize()
    return members, keep, Ls


def batch_shells(Ls, seed=synth.MASTER_SEED + 13):
    """Shell table of the batched form (bins 2..70 generated), the S(k)
    range (wn 0..59: bins 2..61) and per trajectory the three Fs shells
    bracketing the first peak, bins b*-1..b*+1 with b* = round(7.25 L/pi)."""
    off, nv = _abi.shells_generate(2, 69, 300, seed=seed)
    sk_last = 59
    bstar = np.rint(7.25 * np.array(Ls) / np.pi).astype(np.int32)
    fs_first = (bstar - 1 - 2).astype(np.int32)
    fs_last = (bstar + 1 - 2).astype(np.int32)
    return off, nv, sk_last, fs_first, fs_last


def member_traj(cols, L):
    """synth.Traj of a device member (for the CPU oracle)."""
    x, y, z, ix, iy, iz = [t.cpu().numpy() for t in cols]
    T = x.shape[0]
    t = synth.Traj(x, y, z, ix, iy, iz, np.full((T, 3), np.float32(L), np.float32))
    t.L = float(np.float32(L))
    return t
