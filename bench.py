"""Benchmark of the reciprocal-space / time-correlation path on one MI355X.

Workload (precision form): bead-spring melt of 400 chains x 160 beads
(N = 64,000), cubic box L = 42.2, blocked-exponential time scheme
`exponential 10 100 1.2 0 0 1` (T = 1,001 frames, 110 time gaps,
1,046 time-origin pairs with fullblock = 0), Gaussian steps sigma = 0.02.
One step is the whole path:
  rho_k(t) for every frame on 100 shells (bins 2,4,...,200 at k = 2 pi n/L,
           |k| up to 14.9; the isf command's box-L grid),
  F(k,t) over all pairs and S(k) from rho,
  Fs(k,t) on 99 shells at max_length_scale 21.05 (bins 2..100, |k| to 14.9),
  MSD and the non-Gaussian parameter's fourth moment.
updates = N*K_F*T (rho/S) + N*K_Fs*P (Fs) + K_F*P (F).

Prints one JSON line (see the keys below).  --gpus > 1 is refused: the path
is measured on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fs/F(k,t)+S(k) atom·k·time-pair updates/s, 1 GPU, ≤1e-6 rel err vs plain FP64"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0
ERROR_METRIC = ("max |y - y64| / max(|y64|, s): s = |S| for S(k), S(k) of the shell for F(k,t), 1 for Fs(k,t) "
                "(|Fs| <= 1, so its figure is the absolute error), |MSD| for the MSD from the first lag on")


def cpu_model():
    """The host CPU's model name (as lscpu reports it, with its family and
    model numbers, which name the core generation where a virtual machine
    shortens the name) and the logical CPUs."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if not line.strip():
                    break  # the first processor's block
                k, _, v = line.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        pass
    name = fields.get("model name", "unknown")
    if "cpu family" in fields and "model" in fields:
        name += " (family %s, model %s)" % (fields["cpu family"], fields["model"])
    return name, os.cpu_count() or 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(small):
    from amdat_amd import workloads
    return workloads.precision(small, log=log)


def cpu_baseline(tr, shF, shS, L, ell_fs, n_sample, pairs):
    """Reference-arithmetic FP32 restatement (oracle) on a sample: rho for
    one frame and Fs for one pair over n_sample atoms, all vectors.  One core
    (the reference's hot path is serial) is the reported value; the same
    restatement on all host cores (threads over vectors and shells, same
    results) and the FP64 oracle on all host cores are timed beside it."""
    from oracle import oracle as orc
    atoms = np.arange(n_sample, dtype=np.int32)
    tg, t0, t1, ntg = pairs
    sel = (np.array([tg[-2]]), np.array([t0[-2]]), np.array([t1[-2]]))
    f = 500 % tr.T
    upd = n_sample * len(shF[1]) + n_sample * len(shS[1])

    def run(threads, fp64=False):
        orc.set_threads(threads)
        t = time.perf_counter()
        if fp64:
            orc.rho_f64(tr.x[f], tr.y[f], tr.z[f], shF[1], float(np.float32(L)), atoms=atoms)
            orc.fs_f64(tr, sel, ntg, shS[0], shS[1], float(np.float32(ell_fs)), atoms=atoms)
        else:
            orc.ref32_rho(tr.x[f], tr.y[f], tr.z[f], shF[1], np.float32(L), atoms=atoms)
            orc.ref32_fs(tr, sel, ntg, shS[0], shS[1], np.float32(ell_fs), atoms=atoms)
        return time.perf_counter() - t

    dt = run(1)
    cores = max(1, min(16, os.cpu_count() or 1))
    dt_all = run(cores)
    dt64 = run(cores, fp64=True)
    orc.set_threads(1)
    model, ncpu = cpu_model()
    return dict(value=upd / dt, unit="updates/s", cores=1, kind="port",
                sample="rho_k one frame + Fs one time-origin pair, %d atoms x %d/%d vectors, "
                       "reference FP32 arithmetic (cosf/sinf per update), %.1f s" % (
                           n_sample, len(shF[1]), len(shS[1]), dt),
                rate="sample-based: the sample's updates over its time, not a timed full workload",
                cpu_model=model, host_logical_cpus=ncpu,
                omits="the reference's per-(atom, shell) copy of the shell's vector list "
                      "(src/wave_density.cpp:352, src/incoherent_scattering_function.cpp:259) and its "
                      "virtual call per atom: the real reference runs slower than this port",
                all_cores={"value": upd / dt_all, "cores": cores, "kind": "port"},
                fp64_all_cores={"value": upd / dt64, "cores": cores, "kind": "FP64 oracle"})


def bench_batch(args):
    """Batched form: 256 trajectories, one persistent launch per step."""
    import torch
    from amdat_amd import _abi, synth, workloads
    nexp, nes = 4, 50
    n_traj = 32 if args.small else 256
    t = time.time()
    ctx = _abi.Context(0)
    members, keep, Ls = workloads.batch_members(ctx, n_traj, 500, 20000, nexp, nes, synth.MASTER_SEED + 2)
    Ns = np.array([m.N for m in members], np.int64)
    T = members[0].T
    log("generated %d trajectories on the device (sum N = %d, T = %d) in %.1fs" % (n_traj, Ns.sum(), T,
                                                                                     time.time() - t))
    off, nv, sk_last, fs_first, fs_last = workloads.batch_shells(Ls)   # S(k) bins 2..61, Fs b*-1..b*+1
    sh = _abi.Shells(ctx, off, nv)
    tg, t0, t1 = _abi.ts_pairs(nexp, nes, 0, 0, 0)
    ntg = _abi.ts_n_timegaps(nexp, nes, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    b = _abi.Batch(ctx, members)
    K_S = int(off[sk_last + 1])
    K_fs = (off[fs_last + 1] - off[fs_first]).astype(np.int64)
    U_S = int(Ns.sum()) * K_S * T
    U_Fs = int(np.sum(Ns * K_fs)) * len(tg)
    U_msd = int(Ns.sum()) * len(tg)
    updates = U_S + U_Fs
    log("batch: S(k) %d vectors x %d frames, Fs %d..%d vectors x %d pairs; updates/step %.3e"
        % (K_S, T, K_fs.min(), K_fs.max(), len(tg), updates))
    run = lambda: b.run(Ls, sk=(sh, 0, sk_last), fs=(sh, fs_first, fs_last), pairs=P, msd=True,  # noqa: E731
                        graph=True, tune=not args.no_tune)
    t = time.time()
    for i in range(max(1, args.warmup)):
        run()
        log("warmup %d done (%.1fs; plan, tuning and capture on the first)" % (i, time.time() - t))
    info = b.job_settings()
    log("job settings: %s" % info)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        out = run()
        log("step %d done" % i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms = 1e3 * wall / args.steps
    value = updates * args.steps / wall
    tf = (8.0 * U_S + 4.0 * U_Fs) / (ms * 1e-3) / 1e12
    S = out["S"]
    fp64 = None
    if not args.no_fp64:
        # The same job on the plain-FP64 path, one trajectory at a time (FP64
        # phase, FP64 sincos, FP64 sums): its wall time and the batched
        # outputs' error against it.
        torch.cuda.synchronize()
        t = time.perf_counter()
        S64 = np.zeros_like(S)
        Fs64 = []
        for i, m in enumerate(members):
            r = _abi.Rho(ctx, m, None, sh, 0, sk_last, np.float32(Ls[i]), precision=_abi.PREC_FP64)
            S64[i] = _abi.sk(ctx, r)
            r.close()
            f, _ = _abi.fs(ctx, m, None, sh, int(fs_first[i]), int(fs_last[i]), np.float32(Ls[i]), P,
                           precision=_abi.PREC_FP64)
            Fs64.append(f)
            _abi.msd_ngp(ctx, m, None, P)
        torch.cuda.synchronize()
        ms64 = 1e3 * (time.perf_counter() - t)
        err_S = float(np.max(np.abs(S - S64) / np.abs(S64)))
        err_Fs = float(max(np.max(np.abs(a - b)) for a, b in zip(out["Fs"], Fs64)))
        fp64 = {"ms_per_step": ms64, "speedup": ms64 / ms, "rel_err": {"S": err_S, "Fs": err_Fs},
                "within_gate": bool(err_S <= 1e-6 and err_Fs <= 1e-6)}
        log("plain FP64, one trajectory at a time: %.1f ms (%.1fx); errors S %.2e Fs %.2e"
            % (ms64, ms64 / ms, err_S, err_Fs))
    result = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched form: %d trajectories, N uniform in [500, 20000] (sum %d), density "
                               "U[1.0,1.2], exponential %d %d 1.25 (T=%d), %d pairs (fullblock=0); S(k) on 60 "
                               "shells (%d vectors, all frames), Fs on 3 shells at the first peak per "
                               "trajectory, MSD/NGP; one persistent launch per step (graph replay)"
                               % (n_traj, Ns.sum(), nexp, nes, T, len(tg), K_S),
                   "n_trajectories": n_traj, "sum_atoms": int(Ns.sum()), "n_frames": T, "n_pairs": len(tg),
                   "k_vectors_S": K_S, "msd_atom_pairs": U_msd},
        "roofline": {"bound": "valu", "kernel": "job_kernel", "achieved": tf, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS, "traffic": None},
        "job": info,
        "check": {"S_mean_high_k": float(np.mean(S[:, -10:]))},
    }
    if fp64 is not None:
        result["fp64_ms_per_step"] = fp64["ms_per_step"]
        result["speedup_vs_fp64"] = fp64["speedup"]
        result["rel_err_vs_fp64"] = fp64["rel_err"]
        result["error_metric"] = ERROR_METRIC
        result["within_gate"] = fp64["within_gate"]
    print(json.dumps(result), flush=True)


def bench_scale(args):
    """Scale form: N = 1,048,576 atoms, L = 96, `exponential 20 100 1.15`
    (T = 2,001), all time-origin pairs (fullblock = 1, 21,001 pairs); rho,
    S(k) and F(k,t) on 120 shells (bins 2..121), Fs on 5 shells (bins 20,
    40, 60, 80, 100), MSD/NGP.  Positions uniform at frame 0, Gaussian steps
    sigma = 0.05, generated on the device frame by frame (37.8 GB resident)."""
    import torch
    from amdat_amd import _abi, synth
    N, L, nexp, nes, sigma = 1048576, 96.0, 20, 100, 0.05
    if args.small:
        N, nexp = 65536, 4
    T = nes * nexp + 1
    t = time.time()
    ctx = _abi.Context(0)
    cols = synth.device_random_walk(N, L, T, sigma, synth.MASTER_SEED + 4)
    box = np.full((T, 3), L, np.float32)
    d = _abi.Trajectory.from_device(ctx, N, T, [c.data_ptr() for c in cols], box)
    log("generated N=%d T=%d on the device in %.1fs (%.1f GB)" % (N, T, time.time() - t, N * T * 18 / 1e9))
    offA, nA = _abi.shells_generate(2, 120, 300, seed=synth.MASTER_SEED + 14)        # bins 2..121
    shF = _abi.Shells(ctx, offA, nA)
    fs_bins = [20, 40, 60, 80, 100]
    rows = [nA[offA[b - 2]:offA[b - 1]] for b in fs_bins]
    offS = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    nS = np.concatenate(rows).astype(np.int32)
    shS = _abi.Shells(ctx, offS, nS)
    tg, t0, t1 = _abi.ts_pairs(nexp, nes, 0, 1, 0)
    ntg = _abi.ts_n_timegaps(nexp, nes, 0)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    KF, KS, NP = len(nA), len(nS), len(tg)
    U_rho, U_fs, U_f = N * KF * T, N * KS * NP, KF * NP
    updates = U_rho + U_fs + U_f
    log("scale: K_F=%d (120 shells) K_Fs=%d (5 shells) pairs=%d updates/step=%.3e" % (KF, KS, NP, updates))
    Lf = np.float32(L)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    tr0 = time.perf_counter()
    rho = _abi.Rho(ctx, d, None, shF, 0, 119, Lf)
    t_rho = time.perf_counter() - tr0
    log("rho %.1f s" % t_rho)
    F, cnt = _abi.fkt(ctx, rho, rho, P)
    S = _abi.sk(ctx, rho)
    tf0 = time.perf_counter()
    Fs, _ = _abi.fs(ctx, d, None, shS, 0, 4, Lf, P)
    t_fs = time.perf_counter() - tf0
    log("Fs %.1f s" % t_fs)
    s2, s4, c = _abi.msd_ngp(ctx, d, None, P)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # F(k,0) equals S(k) over the same time origins (the lag-0 pairs'
    # frames), to the compensated sums' rounding (the density over fewer
    # frames is split over atoms differently)
    origins = np.unique(t0[tg == 0])
    r0 = _abi.Rho(ctx, d, None, shF, 0, 119, Lf, frames=origins)
    S0 = _abi.sk(ctx, r0)
    r0.close()
    checks = {"Fs(k,0)-1": float(np.max(np.abs(Fs[0] - 1.0))),
              "F(k,0)/S(k)-1 (same origins)": float(np.max(np.abs(F[0] / S0 - 1.0))),
              "MSD(0)": float(s2[0]),
              "S(k) mean high-k": float(np.mean(S[-20:]))}
    tf = 8.0 * U_rho / t_rho / 1e12
    result = {
        "metric": METRIC, "value": updates / wall, "unit": "updates/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "scale form: N=%d, L=%.0f, exponential %d %d 1.15 (T=%d), all time-origin pairs "
                               "(fullblock=1, %d pairs); rho/S/F on 120 shells (%d vectors), Fs on 5 shells "
                               "(%d vectors); MSD/NGP" % (N, L, nexp, nes, T, NP, KF, KS),
                   "n_atoms": N, "n_frames": T, "n_pairs": NP, "resident_gb": N * T * 18 / 1e9},
        "roofline": {"bound": "valu", "kernel": "rho_c32_kernel (incl. host call overhead)", "achieved": tf,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS, "traffic": None},
        "seconds": {"rho": t_rho, "fs": t_fs, "total": wall},
        "invariants": checks,
    }
    print(json.dumps(result), flush=True)


def bench_adjoint(args):
    """Differentiation form: the S(k)-misfit loss and its gradient with
    respect to every coordinate of one frame (amdat_sk_loss_grad: forward
    rho with exact phases, then the adjoint), N = 32,768 sites of a jittered
    fcc lattice (21^3 cells, seeded subset), L = 30.3, 2,000 vectors in 50
    shells (bins 47..96, 40 vectors each), target S(k) from a second seeded
    configuration.  The error is against a plain-FP64 torch evaluation of
    the same loss and its autograd gradient on the device."""
    import torch
    from amdat_amd import _abi, synth
    N, cells, L = 32768, 21, float(np.float32(30.3))
    a = synth.fcc_lattice(cells, L, jitter=0.1, n_keep=N, seed=synth.MASTER_SEED + 3)
    b = synth.fcc_lattice(cells, L, jitter=0.2, n_keep=N, seed=synth.MASTER_SEED + 103)
    off, nvec = _abi.shells_generate(47, 50, 40, seed=synth.MASTER_SEED + 13)
    K = len(nvec)
    ctx = _abi.Context(0)
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    ctx.set_stream(stream.cuda_stream)
    sh = _abi.Shells(ctx, off, nvec)
    pa = [torch.tensor(c[0], device="cuda:0") for c in (a.x, a.y, a.z)]
    pb = [torch.tensor(c[0], device="cuda:0") for c in (b.x, b.y, b.z)]
    w = np.ones(50)
    _, target = _abi.sk_loss_grad(ctx, N, *[t.data_ptr() for t in pb], sh, 0, 49, np.float32(L), w, np.zeros(50))
    grad = [torch.zeros(N, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    ptrs = [t.data_ptr() for t in pa]
    gptrs = [g.data_ptr() for g in grad]

    def step():
        return _abi.sk_loss_grad(ctx, N, *ptrs, sh, 0, 49, np.float32(L), w, target, gptrs)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        loss, S = step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ev_ms = e0.elapsed_time(e1) / args.steps
    # plain FP64 on the device: torch complex sums and autograd
    kv = torch.tensor(nvec, dtype=torch.float64, device="cuda:0") * (2.0 * np.pi / L)
    P = torch.stack([t.double() for t in pa], 1).requires_grad_(True)
    ph = P @ kv.T
    rho2 = torch.cos(ph).sum(0) ** 2 + torch.sin(ph).sum(0) ** 2
    shell = torch.repeat_interleave(torch.arange(50, device="cuda:0"), torch.tensor(np.diff(off), device="cuda:0"))
    S64 = torch.zeros(50, dtype=torch.float64, device="cuda:0").index_add_(0, shell, rho2)
    S64 = S64 / (torch.tensor(np.diff(off), dtype=torch.float64, device="cuda:0") * N)
    loss64 = ((S64 - torch.tensor(target, device="cuda:0")) ** 2).sum()
    loss64.backward()
    g64 = P.grad.detach()
    g = torch.stack(grad, 1)
    errs = {"loss": abs(loss - loss64.item()) / abs(loss64.item()),
            "S": float(np.max(np.abs(S - S64.detach().cpu().numpy()) / S64.detach().cpu().numpy())),
            "grad": float((g - g64).abs().max() / g64.abs().max())}
    del ph, rho2, P
    U = 2 * N * K                      # forward phase updates + adjoint updates
    tf = 8.0 * U / (ev_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        orc.set_threads(1)
        n_s = 8192
        pos = np.stack([a.x[0][:n_s], a.y[0][:n_s], a.z[0][:n_s]], 1).astype(np.float64)
        c0 = time.perf_counter()
        orc.sk_loss_grad_f64(pos, off, nvec, L, w, target)
        dt = time.perf_counter() - c0
        cpu = {"value": 2 * n_s * K / dt, "unit": "updates/s", "cores": 1, "kind": "port",
               "sample": "FP64 oracle loss + gradient, %d sites x %d vectors, %.1f s" % (n_s, K, dt)}
    result = {
        "metric": METRIC, "value": U / wall, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "differentiation form: S(k)-misfit loss + gradient, N=%d jittered fcc sites "
                               "(21^3 cells), L=30.3, %d vectors in 50 shells (bins 47..96)" % (N, K),
                   "n_atoms": N, "k_vectors": K},
        "roofline": {"bound": "valu", "kernel": "rho_direct_kernel + sk_adjoint_kernel (one call)",
                     "achieved": tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS,
                     "traffic": None, "call_ms": ev_ms},
        "rel_err_vs_fp64": errs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def bench_extras(args):
    """The pair-loop statistics, selections and bins on the 64,000-atom melt
    (T = 1,001, 1,046 pairs): per entry point the call time (synchronous
    calls, results copied to the host) and the algorithmic HBM bytes of its
    kernels over that time.  Kernel durations: rocprofv3 --kernel-trace on
    this workload."""
    from amdat_amd import _abi
    cfg, tr, _, _, (tg, t0, t1, ntg) = build_workload(args.small)
    ctx = _abi.Context(0)
    d = _abi.Trajectory(ctx, *tr.arrays)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    N, T, NP = tr.N, tr.T, len(tg)
    nexp, nes = cfg["nexp"], cfg["nes"]
    gtg, g0, g1 = _abi.ts_gap_pairs(nexp, nes, 0, 0, 10)
    G = _abi.Pairs(ctx, gtg, g0, g1, ntg)
    lo = np.zeros((T, 3), np.float32)
    hi = np.full((T, 3), np.float32(cfg["L"]), np.float32)
    cluster = _abi.AtomList(ctx, d, np.arange(0, N, 64))           # 1,000 atoms
    Lh = float(np.float32(cfg["L"]) / 2)
    # algorithmic bytes: 18 B per distinct atom-frame the pairs touch
    # (FP32 xyz + int16 images), 16 B per atom-frame for the bins
    fr_all = len(np.unique(np.concatenate([t0, t1])))
    fr_gap = len(np.unique(np.concatenate([g0, g1])))
    # persistence form of 2x2x2 bins: one key per distinct frame pair
    keys = sorted(set(zip(t0.tolist(), t1.tolist())))
    k0 = np.array([k[0] for k in keys], np.int32)
    k1 = np.array([k[1] for k in keys], np.int32)
    bin_cells = _abi.Cells.assign(ctx, d, (2, 2, 2), lo, hi)
    bin_pl = bin_cells.pair_lists(k0, k1)
    calls = {
        "msd_ngp": (lambda: _abi.msd_ngp(ctx, d, None, P), 18.0 * N * fr_all),
        "displacements (100 van Hove bins)": (lambda: _abi.displacements(ctx, d, None, P, 100, 3.0),
                                              18.0 * N * fr_all),
        "displacement_dist (one gap, 10 pairs)": (lambda: _abi.displacement_dist(ctx, d, None, G, 100, 4.0),
                                                  18.0 * N * fr_gap),
        "select_displacement (find_fast, one gap)": (
            lambda: _abi.TimedList.select_displacement(ctx, d, None, nexp, nes, 0, 0, 10, 0, 0.5, 0.0).close(),
            2 * 18.0 * N * fr_gap),
        "bins_assign (10x10x10)": (lambda: _abi.bins_assign(ctx, d, (10, 10, 10), lo, hi), 16.0 * N * T),
        "bins_distance point": (lambda: _abi.bins_distance(ctx, d, None, _abi.DIST_POINT, 0.5, 40,
                                                           point=(Lh, Lh, Lh)), 16.0 * N * T),
        "bins_distance list (1,000 atoms)": (lambda: _abi.bins_distance(ctx, d, None, _abi.DIST_LIST, 0.25, 16,
                                                                        cluster=cluster), 16.0 * N * T),
        "cells_assign + bin_lists (2x2x2, device)": (
            lambda: [b.close() for b in _abi.Cells.assign(ctx, d, (2, 2, 2), lo, hi).bin_lists()], 16.0 * N * T),
        "cells_assign + pair_lists (2x2x2, persistence keys)": (
            lambda: [b.close() for b in _abi.Cells.assign(ctx, d, (2, 2, 2), lo, hi).pair_lists(k0, k1)],
            8.0 * N * len(k0) + 4.0 * N * T),
        "msd_ngp on one pair-keyed bin list (one launch)": (lambda: _abi.msd_ngp(ctx, d, bin_pl[0], P),
                                                             18.0 * N * fr_all / 8),
    }
    res = {}
    for name, (fn, nbytes) in calls.items():
        fn()
        ts = []
        for _ in range(max(1, args.steps)):
            c0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - c0)
        t = float(np.median(ts))
        res[name] = {"call_ms": 1e3 * t, "algorithmic_bytes": nbytes, "GBps_at_call_time": nbytes / t / 1e9}
        log("%-42s %9.3f ms  %8.1f GB/s" % (name, 1e3 * t, nbytes / t / 1e9))
    result = {"workload": "pair-loop statistics, selections and bins on the precision-form melt (N=%d, T=%d, "
                          "%d pairs)" % (N, T, NP), "calls": res, "peak_hbm_GBps": PEAK_HBM_GBS}
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="reduced workload for a quick check")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain-FP64 error check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--workload", choices=["precision", "batch", "scale", "adjoint", "extras"], default="precision",
                    help="precision: the 64,000-atom melt (default); batch: 256 trajectories in one launch; "
                         "scale: 1,048,576 atoms x 2,001 frames, all time-origin pairs; adjoint: S(k)-misfit "
                         "loss and gradient, 32,768 atoms; extras: pair-loop statistics, selections and bins")
    ap.add_argument("--no-tune", action="store_true", help="batch workload: skip the autotuner")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the path is built and measured on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "scale":
        bench_scale(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "extras":
        bench_extras(args)
        return

    import torch
    from amdat_amd import _abi

    cfg, tr, (offF, nF), (offS, nS), (tg, t0, t1, ntg) = build_workload(args.small)
    L = np.float32(cfg["L"])
    ell_fs = np.float32(cfg["ell_fs"])
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = _abi.Context(0)
    ctx.set_stream(stream.cuda_stream)
    d = _abi.Trajectory(ctx, *tr.arrays)
    shF = _abi.Shells(ctx, offF, nF)
    shS = _abi.Shells(ctx, offS, nS)
    P = _abi.Pairs(ctx, tg, t0, t1, ntg)
    nshF, nshS = len(offF) - 1, len(offS) - 1
    KF, KS, NP, N, T = len(nF), len(nS), len(tg), tr.N, tr.T
    U_rho = N * KF * T
    U_fs = N * KS * NP
    U_f = KF * NP
    updates = U_rho + U_fs + U_f
    log("workload: N=%d T=%d K_F=%d (%d shells) K_Fs=%d (%d shells) pairs=%d updates/step=%.3e"
        % (N, T, KF, nshF, KS, nshS, NP, updates))

    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(prec, timings=None):
        e0, e1, e2, e3 = ev(), ev(), ev(), ev()
        e0.record(stream)
        rho = _abi.Rho(ctx, d, None, shF, 0, nshF - 1, L, precision=prec)
        e1.record(stream)
        F, cnt = _abi.fkt(ctx, rho, rho, P)
        S = _abi.sk(ctx, rho)
        e2.record(stream)
        Fs, _ = _abi.fs(ctx, d, None, shS, 0, nshS - 1, ell_fs, P, precision=prec)
        e3.record(stream)
        s2, s4, c = _abi.msd_ngp(ctx, d, None, P)
        rho.close()
        if timings is not None:
            torch.cuda.synchronize()
            timings["rho"].append(e0.elapsed_time(e1))
            timings["fs"].append(e2.elapsed_time(e3))
        return dict(S=S, F=F, Fs=Fs, msd=s2 / np.maximum(c, 1), s4=s4, count=c)

    for i in range(args.warmup):
        step(_abi.PREC_FP32C)
        log("warmup %d done" % i)
    timings = {"rho": [], "fs": []}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        out = step(_abi.PREC_FP32C, timings)
        log("step %d: rho %.1f ms, Fs %.1f ms" % (i, timings["rho"][-1], timings["fs"][-1]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * wall / args.steps
    value = updates * args.steps / wall

    result = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "precision form: bead-spring melt %d chains x %d beads (N=%d), L=%.1f, "
                               "exponential %d %d %.2f (T=%d), %d pairs (fullblock=0); rho/F/S on %d shells "
                               "(%d vectors, box-L grid), Fs on %d shells (%d vectors, l=%.2f); MSD/NGP"
                               % (cfg["n_chains"], cfg["beads"], N, cfg["L"], cfg["nexp"], cfg["nes"], cfg["base"],
                                  T, NP, nshF, KF, nshS, KS, cfg["ell_fs"]),
                   "n_atoms": N, "n_frames": T, "n_pairs": NP, "k_vectors_F": KF, "k_vectors_Fs": KS},
    }

    # Roofline of the dominant kernel (FP32 VALU bound): algorithmic flops per
    # launch / measured launch time.  rho: 8 flops per (atom, vector, frame);
    # Fs: 4 flops per (atom, vector, pair).
    rho_ms = float(np.mean(timings["rho"]))
    fs_ms = float(np.mean(timings["fs"]))
    rho_tf = 8.0 * U_rho / (rho_ms * 1e-3) / 1e12
    fs_tf = 4.0 * U_fs / (fs_ms * 1e-3) / 1e12
    dom = "rho" if rho_ms >= fs_ms else "fs"
    ach = rho_tf if dom == "rho" else fs_tf
    kname = "rho_c32_kernel" if dom == "rho" else "fs_c32_kernel"
    result["roofline"] = {"bound": "valu", "kernel": kname,
                          "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                          "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                          "launch_ms": {"rho": rho_ms, "fs": fs_ms},
                          "tflops": {"rho": rho_tf, "fs": fs_tf}}
    # HBM bytes per launch of the same kernel on this workload, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (counters
    # cannot be read from inside the timed run).
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not args.small and os.path.exists(tpath):
        t = json.load(open(tpath))
        k = t["kernels"].get(kname)
        if k:
            result["roofline"]["traffic"] = int(1024 * (k["fetch_kib"] * t["fetch_factor"]
                                                        + k["write_kib"] * t["write_factor"]))
            result["roofline"]["traffic_source"] = "profiles/hbm_traffic.json: %s" % k["instance"]

    if not args.no_fp64:
        t = time.perf_counter()
        ref = step(_abi.PREC_FP64)
        torch.cuda.synchronize()
        fp64_ms = 1e3 * (time.perf_counter() - t)
        S64 = ref["S"]
        err = {
            "S": float(np.max(np.abs(out["S"] - S64) / np.abs(S64))),
            "F": float(np.max(np.abs(out["F"] - ref["F"]) / np.maximum(np.abs(ref["F"]), np.abs(S64)[None]))),
            "Fs": float(np.max(np.abs(out["Fs"] - ref["Fs"]) / np.maximum(np.abs(ref["Fs"]), 1.0))),
            "MSD": float(np.max(np.abs(out["msd"][1:] - ref["msd"][1:]) / np.abs(ref["msd"][1:]))),
        }
        result["rel_err_vs_fp64"] = err
        result["error_metric"] = ERROR_METRIC
        result["within_gate"] = bool(all(v <= 1e-6 for v in err.values()))
        result["fp64_ms_per_step"] = fp64_ms
        result["speedup_vs_fp64"] = fp64_ms / ms_per_step
        log("plain FP64 step %.1f ms; errors %s" % (fp64_ms, err))
        # The tabled-FP64 form: the same plan and phase tables as the timed
        # FP32 form with FP64 entries, products and sums (the strongest FP64
        # form here); one untimed run first, then one timed step.
        step(_abi.PREC_FP64T)
        t64 = {"rho": [], "fs": []}
        torch.cuda.synchronize()
        t = time.perf_counter()
        tab = step(_abi.PREC_FP64T, t64)
        torch.cuda.synchronize()
        tab_ms = 1e3 * (time.perf_counter() - t)
        tab_err = {
            "S": float(np.max(np.abs(tab["S"] - S64) / np.abs(S64))),
            "F": float(np.max(np.abs(tab["F"] - ref["F"]) / np.maximum(np.abs(ref["F"]), np.abs(S64)[None]))),
            "Fs": float(np.max(np.abs(tab["Fs"] - ref["Fs"]) / np.maximum(np.abs(ref["Fs"]), 1.0))),
        }
        result["fp64_tabled_ms_per_step"] = tab_ms
        result["speedup_vs_fp64_tabled"] = tab_ms / ms_per_step
        result["fp64_tabled_launch_ms"] = {"rho": t64["rho"][0], "fs": t64["fs"][0]}
        result["fp64_tabled_rel_err_vs_fp64"] = tab_err
        log("tabled FP64 step %.1f ms (rho %.1f ms, Fs %.1f ms): FP32 form %.2fx faster; vs plain FP64 %s"
            % (tab_ms, t64["rho"][0], t64["fs"][0], tab_ms / ms_per_step, tab_err))

    if not args.no_cpu:
        n_sample = 600 if args.small else 20000   # about 10 s of one core
        result["cpu_baseline"] = cpu_baseline(tr, (offF, nF), (offS, nS), cfg["L"], cfg["ell_fs"], n_sample,
                                              (tg, t0, t1, ntg))
        cb = result["cpu_baseline"]
        log("cpu baseline %.3e updates/s (1 core); %d cores: %.3e; FP64 oracle on %d cores: %.3e"
            % (cb["value"], cb["all_cores"]["cores"], cb["all_cores"]["value"], cb["fp64_all_cores"]["cores"],
               cb["fp64_all_cores"]["value"]))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
