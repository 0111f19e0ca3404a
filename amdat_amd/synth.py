"""Seeded synthetic trajectories for the benchmark configurations.

All generators build unwrapped positions in FP64 and store what a LAMMPS
``custom`` dump with ``x y z ix iy iz`` would hold: wrapped FP32
coordinates ``u - L*floor(u/L)`` and int16 image counts ``floor(u/L)``,
frame-major SoA arrays of shape [T, N].  One Gaussian step is taken per
stored frame.  Random numbers come from numpy's PCG64 seeded by
``0xA11DA7 + config index`` unless a seed is given.
"""
import numpy as np

MASTER_SEED = 0xA11DA7


class Traj:
    """Host trajectory in the device layout."""

    def __init__(self, x, y, z, ix, iy, iz, box, types=None):
        self.x, self.y, self.z = x, y, z
        self.ix, self.iy, self.iz = ix, iy, iz
        self.box = box
        self.types = types
        self.T, self.N = x.shape

    @property
    def arrays(self):
        return (self.x, self.y, self.z, self.ix, self.iy, self.iz, self.box)

    def unwrapped64(self):
        """FP64 unwrapped positions [T, N, 3] (x + ix*L)."""
        L = self.box.astype(np.float64)
        u = np.empty((self.T, self.N, 3))
        u[..., 0] = self.x + self.ix * L[:, 0:1]
        u[..., 1] = self.y + self.iy * L[:, 1:2]
        u[..., 2] = self.z + self.iz * L[:, 2:3]
        return u


def _store(u, L):
    """u: FP64 unwrapped [T, N, 3] (may be built frame by frame)."""
    img = np.floor(u / L)
    w = (u - L * img).astype(np.float32)
    # FP32 rounding can land a coordinate exactly on L; keep the image
    # count consistent so x + ix*L is still the unwrapped position.
    hit = w >= np.float32(L)
    if np.any(hit):
        w[hit] = 0.0
        img[hit] += 1
    if np.abs(img).max(initial=0) > 32767:
        raise ValueError("image counts exceed int16")
    img = img.astype(np.int16)
    return [np.ascontiguousarray(w[..., c]) for c in range(3)], [np.ascontiguousarray(img[..., c]) for c in range(3)]


def random_walk(N, L, T, sigma, seed, start=None):
    """Independent Gaussian steps of std sigma per axis per frame."""
    rng = np.random.default_rng(seed)
    u = np.empty((T, N, 3))
    u[0] = rng.uniform(0.0, L, size=(N, 3)) if start is None else start
    steps = rng.normal(0.0, sigma, size=(T - 1, N, 3))
    np.cumsum(steps, axis=0, out=steps)
    u[1:] = u[0][None] + steps
    (x, y, z), (ix, iy, iz) = _store(u, L)
    box = np.full((T, 3), L, np.float32)
    return Traj(x, y, z, ix, iy, iz, box)


def random_walk_box(N, box, T, sigma, seed):
    """random_walk in an orthorhombic box (Lx, Ly, Lz)."""
    rng = np.random.default_rng(seed)
    L = np.asarray(box, np.float64)
    u = np.empty((T, N, 3))
    u[0] = rng.uniform(0.0, 1.0, size=(N, 3)) * L
    steps = rng.normal(0.0, sigma, size=(T - 1, N, 3))
    np.cumsum(steps, axis=0, out=steps)
    u[1:] = u[0][None] + steps
    (x, y, z), (ix, iy, iz) = _store(u, L)
    return Traj(x, y, z, ix, iy, iz, np.tile(L.astype(np.float32), (T, 1)))


def plain_config(N=1000, L=9.41, nexp=5, nes=20, sigma=0.05, seed=MASTER_SEED + 0):
    """3D binary 80:20 mixture, uniform frame 0, Gaussian steps."""
    T = nes * nexp + 1
    t = random_walk(N, L, T, sigma, seed)
    t.types = np.where(np.arange(N) < int(0.8 * N), 1, 2).astype(np.int32)
    return t


def batch_configs(n_traj, n_min, n_max, nexp, nes, seed=MASTER_SEED + 2):
    """Many independent systems of uneven size (the batched workload):
    N_i ~ U{n_min..n_max}, density rho_i ~ U[1.0, 1.2], L_i = (N_i/rho_i)^(1/3),
    T = nes*nexp + 1 frames, Gaussian steps sigma_i ~ U[0.01, 0.1].
    Returns a list of Traj, each with attribute L (its box length)."""
    rng = np.random.default_rng(seed)
    T = nes * nexp + 1
    out = []
    for i in range(n_traj):
        N = int(rng.integers(n_min, n_max + 1))
        dens = rng.uniform(1.0, 1.2)
        L = float(np.float32((N / dens) ** (1.0 / 3.0)))
        sigma = rng.uniform(0.01, 0.1)
        t = random_walk(N, L, T, sigma, seed=int(rng.integers(1 << 62)))
        t.L = L
        out.append(t)
    return out


def chain_melt(n_chains=400, beads=160, L=42.2, T=1001, sigma=0.02, seed=MASTER_SEED + 1):
    """Bead-spring melt: seeded random-walk chains (bond U[0.95,0.99],
    uniform directions, uniform chain starts), then Gaussian steps."""
    rng = np.random.default_rng(seed)
    N = n_chains * beads
    bond = rng.uniform(0.95, 0.99, size=(n_chains, beads - 1))
    d = rng.normal(size=(n_chains, beads - 1, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    pos = np.zeros((n_chains, beads, 3))
    pos[:, 0] = rng.uniform(0.0, L, size=(n_chains, 3))
    pos[:, 1:] = pos[:, :1] + np.cumsum(d * bond[..., None], axis=1)
    start = pos.reshape(N, 3)
    xs, ys, zs, ixs, iys, izs = ([np.empty((T, N), np.float32) for _ in range(3)]
                                 + [np.empty((T, N), np.int16) for _ in range(3)])
    cur = start.copy()
    for t in range(T):
        if t > 0:
            cur += rng.normal(0.0, sigma, size=(N, 3))
        (x, y, z), (ix, iy, iz) = _store(cur[None], L)
        xs[t], ys[t], zs[t] = x[0], y[0], z[0]
        ixs[t], iys[t], izs[t] = ix[0], iy[0], iz[0]
    box = np.full((T, 3), L, np.float32)
    return Traj(xs, ys, zs, ixs, iys, izs, box)


def write_lammps_custom(tr, path, types=None, box_lo=0.0):
    """Write the trajectory as a LAMMPS custom dump with columns
    ``id type x y z ix iy iz``; coordinates in %.9g so FP32 values
    round-trip exactly."""
    types = tr.types if types is None else types
    if types is None:
        types = np.ones(tr.N, np.int32)
    with open(path, "w") as fh:
        for t in range(tr.T):
            fh.write("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (t, tr.N))
            fh.write("ITEM: BOX BOUNDS pp pp pp\n")
            for c in range(3):
                fh.write("%.9g %.9g\n" % (box_lo, box_lo + float(tr.box[t, c])))
            fh.write("ITEM: ATOMS id type x y z ix iy iz\n")
            rows = zip(range(1, tr.N + 1), types, tr.x[t], tr.y[t], tr.z[t], tr.ix[t], tr.iy[t], tr.iz[t])
            fh.write("".join("%d %d %.9g %.9g %.9g %d %d %d\n" % r for r in rows))


def fcc_lattice(cells, L, jitter=0.0, n_keep=None, seed=MASTER_SEED + 3):
    """Single frame of an fcc lattice of cells^3 unit cells (4 sites each)
    in a cubic box of edge L, optional Gaussian jitter and a seeded
    subset of n_keep sites."""
    rng = np.random.default_rng(seed)
    a = L / cells
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij"), -1).reshape(-1, 3)
    sites = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    if n_keep is not None:
        sites = sites[np.sort(rng.choice(len(sites), n_keep, replace=False))]
    if jitter:
        sites = sites + rng.normal(0.0, jitter, size=sites.shape)
    (x, y, z), (ix, iy, iz) = _store(sites[None], L)
    box = np.full((1, 3), L, np.float32)
    return Traj(x, y, z, ix, iy, iz, box)


def device_random_walk(N, L, T, sigma, seed):
    """random_walk generated on the GPU frame by frame with torch: FP64
    positions, stored as frame-major FP32 wrapped coordinates and int16
    images [T, N] (CUDA tensors, x y z ix iy iz) -- for systems too large to
    build on the host.  The stream differs from random_walk's."""
    import torch
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    dev = torch.device("cuda:0")
    cols = [torch.empty((T, N), dtype=torch.float32, device=dev) for _ in range(3)] + \
           [torch.empty((T, N), dtype=torch.int16, device=dev) for _ in range(3)]
    u = torch.empty((N, 3), dtype=torch.float64, device=dev).uniform_(0.0, L, generator=gen)
    step = torch.empty_like(u)
    for f in range(T):
        if f:
            u += step.normal_(0.0, sigma, generator=gen)
        img = torch.floor(u / L)
        w = (u - L * img).to(torch.float32)
        hit = w >= np.float32(L)
        w[hit] = 0.0
        img[hit] += 1.0
        for c in range(3):
            cols[c][f] = w[:, c]
            cols[3 + c][f] = img[:, c].to(torch.int16)
    torch.cuda.synchronize()
    return cols
