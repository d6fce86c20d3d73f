"""Benchmark: warm-started vs plain batched PDHG-TV (Poisson data term) on 512 x 512 XRF slices.

Workload (one GPU's share): a batch of 512 x 512 float32 phantoms of 30 ellipses + 5 Gaussian
blobs (seeded), mu_in = 0.02 (f > 0) + 0.005, mu_out = 0.05 (f > 0), 720 angles over 2 pi, two
detectors at +-90 deg, Poisson noise at a peak of 1e3 counts; batched PDHG with the KL data term
and TV lambda = 1e-2, the exact adjoint as back-projection (with the reference's default
back-projection, which ignores the attenuation maps, PDHG-KL diverges on this workload),
HIP-graph iterations, autotuned persistent grids.

A step is one batched PDHG iteration (fp + bp of every slice + the fused dual, TV and primal
kernels); K steps are timed after W warm-up steps. The iteration counts come from convergence
runs during set-up against x*, a long plain PDHG run whose own drift over its last half is
reported: plain PDHG from 0, and the warm-started method (batched SIRT-20, the U-Net's estimate
of the solution, then PDHG from it; the plain PDHG iteration judges convergence). value = the
warm-started method's slices/s to the tolerance:
    slices / (preconditioner + SIRT-20 + network + its_warm x ms_per_step),
beside the plain method's slices / (preconditioner + its_plain x ms_per_step) and their ratio.

Also reported: the batched fp and bp against the HBM roofline (HIP events on the stream they
run on); the configuration's MLEM-300 on one slice; the oracle's C restatement of the reference
CPU path (same fp and exact-adjoint bp) and a NumPy restatement with the reference's per-angle
thread pool, both timed on bounded samples on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the kernels of one batched two-detector fp per autotuned shape (the tile shapes: a pair of slices
# per workgroup, then the fixed-order sum of the tiles' partial sums)
FP_KERNEL_NAMES = {
    "tile": "fp_tile_kernel<2, 2, true, true, 1> + fp_tile_reduce_kernel<2>",
    "tile_ahead2": "fp_tile_kernel<2, 2, true, true, 2> + fp_tile_reduce_kernel<2>",
    "tile_staged": "fp_tile_kernel<2, 2, true, false, 1> + fp_tile_reduce_kernel<2>",
    "stream": "fp_stream_batch_kernel<2, true, false, false>",
    "stream_trim": "fp_stream_batch_kernel<2, true, true, false>",
    "direct": "fp_direct_batch_kernel<2, false>",
    "direct_trim": "fp_direct_batch_kernel<2, true>",
}
WEIGHTS = os.path.join(ROOT, "pycorrectedemissionct_amd", "data", "unet_pdhg_kl_512x720x2.npz")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def slice_errors(x, ref, vs) -> list[float]:
    import torch

    out = []
    for s in range(len(vs) - 1):
        lo, hi = int(vs[s]), int(vs[s + 1])
        out.append(float(torch.linalg.vector_norm((x[lo:hi] - ref[lo:hi]).double())
                         / torch.linalg.vector_norm(ref[lo:hi].double())))
    return out


def cpu_baseline(maps_np, ang, n: int, its: int, budget_s: float) -> dict:
    """The oracle's C/OpenMP restatement of the reference CPU path: the skimage-semantics fp over
    the precomputed maps and the exact-adjoint bp the GPU path uses, for one slice; fp + bp pairs
    for about ``budget_s`` seconds, scaled to one plain solve (``its`` iterations + preconditioner)."""
    from oracle import cops

    tabs = cops.tables(ang, n)
    vol = np.random.default_rng(0).random((n, n)).astype(np.float32)
    t_fp = t_bp = 0.0
    pairs = 0
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        y = cops.fp(vol, ang, maps_np, tabs=tabs)
        t1 = time.perf_counter()
        cops.bp_adjoint(y, ang, n, maps_np, tabs=tabs)
        t2 = time.perf_counter()
        t_fp += t1 - t0
        t_bp += t2 - t1
        pairs += 1
        if t2 - t_start >= budget_s:
            break
    t_solve = (t_fp + t_bp) / pairs * (its + 1)
    return {
        "value": 1.0 / t_solve, "unit": "slices/s", "cores": int(cops.lib().oracle_num_threads()), "kind": "port",
        "sample": (f"{pairs} fp + exact-adjoint bp pairs of one {n}x{n} slice ({len(ang)} angles x "
                   f"{maps_np.shape[1]} detectors, {t_fp + t_bp:.1f} s) of the oracle's C/OpenMP restatement: "
                   f"fp {1e3 * t_fp / pairs:.0f} ms, bp {1e3 * t_bp / pairs:.0f} ms per call, scaled to the "
                   f"plain method's {its} PDHG iterations + preconditioner (elementwise updates not counted)"),
    }


def cpu_numpy_baseline(maps_np, ang, n: int, its: int, budget_s: float) -> dict:
    """The oracle's NumPy restatement in the reference's execution structure: per-angle fp / bp
    calls on a ThreadPoolExecutor(round(log2(ncpu + 1))) (``projectors.py:20, 585-587``), timed on
    a bounded sample of angles and scaled to all angles and one plain solve."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import xrf as oxrf

    workers = int(round(np.log2((os.cpu_count() or 1) + 1)))
    vol = np.random.default_rng(0).random((n, n)).astype(np.float32)
    line = np.random.default_rng(1).random((maps_np.shape[1], 1, n)).astype(np.float32)
    angles_done = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        while time.perf_counter() - t0 < budget_s and angles_done < len(ang):
            batch = list(range(angles_done, min(angles_done + workers, len(ang))))
            list(ex.map(lambda a: (oxrf.fp(vol, ang[a:a + 1], maps_np[a:a + 1]),
                                   oxrf.bp_skimage(line, ang[a:a + 1], n, maps_np[a:a + 1], symmetric=True)), batch))
            angles_done += len(batch)
    t = time.perf_counter() - t0
    t_pair = t / angles_done * len(ang)
    return {
        "value": 1.0 / (t_pair * (its + 1)), "unit": "slices/s", "threads": workers,
        "sample": (f"{angles_done} of {len(ang)} angles (fp and map-weighted bp per angle and detector) in "
                   f"{t:.1f} s on {workers} threads, scaled to all angles: {t_pair:.1f} s per fp + bp pair, "
                   f"{its} PDHG iterations + preconditioner"),
    }


def run_headline(args) -> dict:
    import torch

    import pycorrectedemissionct_amd as pc
    from pycorrectedemissionct_amd import batch as B
    from pycorrectedemissionct_amd import warmstart as ws

    dev = torch.device("cuda", 0)
    n, n_ang, nb = args.n, args.angles, args.slices
    det = [np.pi / 2, -np.pi / 2]
    ang = pc.testing.angles_uniform(n_ang)
    log(f"set-up: {nb} slices of {n}x{n}, {n_ang} angles, 2 detectors")
    fs, mis, mos, wls = [], [], [], []
    for k in range(nb):
        wl = pc.testing.workload_pdhg_512(seed=20260102 + k, n=n, n_angles=n_ang)
        fs.append(wl["f"]), mis.append(wl["mu_in"]), mos.append(wl["mu_out"]), wls.append(wl)
    A = B.BatchedProjectorXRF([n] * nb, ang, mis, mos, angles_detectors_rad=det, bp_mode="adjoint")
    tune = B.autotune(A)
    clean = A.fp(A.pack_vol(fs))
    ss = A.plan.sino_start
    b = torch.cat([torch.as_tensor(pc.testing.poisson_data(clean[int(ss[s]):int(ss[s + 1])].cpu().numpy(), 1e3,
                                                           wls[s]["rng"]), device=dev) for s in range(nb)])
    vs = A.plan.vol_start
    solver = B.BatchedPDHG(data_term="kl", tv_lambda=1e-2, use_graph=True)

    # x*: a long plain run; its drift over the last half says how far from converged it still is
    t0 = time.perf_counter()
    st = solver.prepare(A, b)
    solver.iterate(A, st, args.ref_iters // 2)
    x_half = st["x"].clone()
    solver.iterate(A, st, args.ref_iters - args.ref_iters // 2)
    x_star = st["x"].clone()
    drift = slice_errors(x_half, x_star, vs)
    log(f"x*: {args.ref_iters} plain iterations ({time.perf_counter() - t0:.0f} s); |x_{args.ref_iters // 2} - x*| / "
        f"|x*| = {max(drift):.2e}")

    def its_to_tol(state, cap):
        done = 0
        while done < cap:
            solver.iterate(A, state, args.check_every)
            done += args.check_every
            if max(slice_errors(state["x"], x_star, vs)) <= args.tol:
                return done
        return None

    cap = args.ref_iters - args.ref_iters // 4
    its_plain = its_to_tol(solver.prepare(A, b), cap)
    log(f"plain PDHG: {its_plain} iterations to {args.tol:g}")
    # the warm-started method: SIRT-20 -> network -> PDHG from its estimate
    net = ws.UNet.load(args.weights)
    sirt = B.BatchedSIRT(use_graph=True)
    xs = torch.stack(A.unpack_vol(sirt(A, b, 20)))
    x0 = A.pack_vol(list(ws.warm_start(net, xs)))
    e0 = slice_errors(x0, x_star, vs)
    its_warm = its_to_tol(solver.prepare(A, b, x0=x0), cap)
    log(f"warm start: |x0 - x*| / |x*| = {max(e0):.2e}; {its_warm} iterations to {args.tol:g}")

    # timing: K batched PDHG iterations after W warm-up iterations (HIP-graph replays)
    stream = torch.cuda.current_stream()

    def sync_time(fn) -> float:
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    st = solver.prepare(A, b)
    solver.iterate(A, st, args.warmup)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    solver.iterate(A, st, args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    t_pre = min(sync_time(lambda: solver.prepare(A, b)) for _ in range(3))
    t_sirt = min(sync_time(lambda: sirt(A, b, 20)) for _ in range(3))
    t_net = min(sync_time(lambda: ws.warm_start(net, xs)) for _ in range(3))
    log(f"{args.steps} steps: {ms_per_step:.3f} ms per batched iteration (wall {1e3 * t_wall / args.steps:.3f}); "
        f"preconditioner {1e3 * t_pre:.1f} ms, SIRT-20 {1e3 * t_sirt:.1f} ms, network {1e3 * t_net:.1f} ms")
    t_warm = None if its_warm is None else t_pre + t_sirt + t_net + its_warm * ms_per_step * 1e-3
    t_plain = None if its_plain is None else t_pre + its_plain * ms_per_step * 1e-3
    warm_sps = None if t_warm is None else nb / t_warm
    plain_sps = None if t_plain is None else nb / t_plain

    # batched fp and bp against the roofline (HIP events on the stream they run on)
    x = st["x"]
    y = torch.empty_like(b)
    v = torch.empty_like(x)

    def time_launches(fn, reps) -> float:
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t_fp = time_launches(lambda: A.fp(x, y), args.fp_reps)
    t_bp = time_launches(lambda: A.bp(y, v), args.fp_reps)
    n_det = 2
    fp_bytes = nb * 4.0 * (n * n + n_ang * n_det * n * n + n_ang * n_det * n)
    bp_bytes = nb * 4.0 * (n_ang * n_det * n + n_ang * n_det * n * n + n * n)
    fp_gbs, bp_gbs = fp_bytes / t_fp / 1e9, bp_bytes / t_bp / 1e9
    log(f"batched fp {t_fp * 1e6:.1f} us ({fp_gbs:.0f} GB/s), bp {t_bp * 1e6:.1f} us ({bp_gbs:.0f} GB/s)")
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    t_copy = time_launches(lambda: dst.copy_(src), 10)
    copy_gbs = 2.0 * 4 * src.numel() / t_copy / 1e9
    del src, dst

    # the configuration's second solve on one slice: MLEM, 300 iterations, through the solver API
    prj = pc.ProjectorAttenuationXRF([n, n], ang, att_in=mis[0], att_out=mos[0], angles_detectors_rad=det,
                                     bp_mode="adjoint")
    b0 = b[int(ss[0]):int(ss[1])].reshape(2, n_ang, n).contiguous()
    mlem = pc.MLEM()
    mlem(prj, b0, iterations=5)
    t_mlem = sync_time(lambda: mlem(prj, b0, iterations=300))
    log(f"MLEM 300 iterations, one slice: {t_mlem * 1e3:.1f} ms")

    tile = tune["fp_shape"] in B.TILE_SHAPES
    kernels = {
        "fp": dict(kernel=FP_KERNEL_NAMES[tune["fp_shape"]], t=t_fp, bytes=fp_bytes, gbs=fp_gbs,
                   traffic_file="fp_tile_batch_2x512x720x2_traffic.json" if tile
                   else "fp_batch_2x512x720x2_traffic.json"),
        "bp": dict(kernel="bp_adjoint_slots_kernel<2, true, 2> + sum_parts_acc_kernel"
                   if tune["bp_wg_per_cu"] == "pairs"
                   else "bp_adjoint_batch_kernel<2, true>", t=t_bp, bytes=bp_bytes, gbs=bp_gbs,
                   traffic_file="bp_batch_2x512x720x2_traffic.json"),
    }
    dom = kernels["fp"] if t_fp >= t_bp else kernels["bp"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", dom["traffic_file"])
    if os.path.exists(tfile):
        try:
            with open(tfile) as fh:
                traffic = json.load(fh).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    ratio = None if (warm_sps is None or plain_sps is None) else warm_sps / plain_sps
    result = {
        "metric": "slices/s to 1e-3 rel err, warm-start vs plain PDHG-TV; fp+bp HBM GB/s",
        "value": warm_sps, "unit": "slices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": (f"warm-started (SIRT-20 + U-Net + PDHG) vs plain batched PDHG-KL-TV(1e-2), {nb} x "
                         f"{n}x{n} XRF slices, {n_ang} angles over 2pi, 2 detectors at +-90 deg, Poisson peak 1e3, "
                         f"exact-adjoint bp; tolerance {args.tol:g} rel err vs x* = {args.ref_iters} plain "
                         f"iterations (1e-3 is out of reach: x* itself still moves {max(drift):.1e} over its last "
                         f"{args.ref_iters // 2} iterations, so the tolerance is kept well above that)"),
            "slices": nb, "n": n, "angles": n_ang, "detectors": n_det, "tolerance": args.tol,
            "x_star_iterations": args.ref_iters, "x_star_drift_last_half": drift,
            "plain_iterations": its_plain, "warm_iterations": its_warm, "warm_x0_rel_err": e0,
            "plain_slices_per_s": plain_sps, "warm_slices_per_s": warm_sps, "warm_over_plain": ratio,
            "ms_preconditioner": 1e3 * t_pre, "ms_sirt20": 1e3 * t_sirt, "ms_network": 1e3 * t_net,
            "ms_per_iteration_per_slice": ms_per_step / nb, "mlem300_ms_one_slice": 1e3 * t_mlem,
            "fp_us": t_fp * 1e6, "bp_us": t_bp * 1e6, "fp_plus_bp_gbs": (fp_bytes + bp_bytes) / (t_fp + t_bp) / 1e9,
            "autotune": {"fp_shape": tune["fp_shape"], "fp_wg_per_cu": tune["fp_wg_per_cu"],
                         "bp_wg_per_cu": tune["bp_wg_per_cu"]},
            "network": os.path.relpath(args.weights, ROOT),
        },
        "roofline": {
            "bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom["kernel"],
            "algorithmic_bytes_per_launch": dom["bytes"], "copy_gbs_measured": copy_gbs,
            "frac_of_copy": dom["gbs"] / copy_gbs,
            "other": {k: {"kernel": v["kernel"], "achieved": v["gbs"], "frac": v["gbs"] / HBM_PEAK_GBS,
                          "us": v["t"] * 1e6} for k, v in kernels.items() if v is not dom},
        },
    }
    if not args.no_cpu and its_plain is not None:
        maps0 = A.slice_maps(0).cpu().numpy()
        log(f"cpu baseline (oracle C restatement, about {args.cpu_seconds:g} s)")
        result["cpu_baseline"] = cpu_baseline(maps0, ang, n, its_plain, args.cpu_seconds)
        log("cpu baseline (NumPy restatement on the reference's per-angle thread pool, about 8 s)")
        result["cpu_baseline"]["numpy"] = cpu_numpy_baseline(maps0, ang, n, its_plain, 8.0)
    return result


def run_batched(args) -> dict:
    """One GPU's share of the batched workload: 32 slices of sides drawn from {128, 256, 384, 512, 768},
    each with its own 10-ellipse phantom and attenuation and 4 emission channels (mu_out scales 0.02,
    0.04, 0.08, 0.16, channel intensities U[0, 1]), 360 angles, one detector at 90 deg, Gaussian noise
    at 1 % of the max; batched SIRT (100 iterations) then PDHG-TV(1e-2, l2; 300 iterations), HIP-graph
    iterations, autotuned persistent grids. ``--maps otf`` rebuilds the maps inside every fp / bp
    instead (the whole 256-slice batch needs about 320 GB of maps); ``--sirt-iters`` /
    ``--pdhg-iters`` shorten the timed solves."""
    import torch

    import pycorrectedemissionct_amd as pc

    rng = np.random.default_rng(20260103)
    sizes_sl = rng.choice([128, 256, 384, 512, 768], size=args.slices)
    scales = [0.02, 0.04, 0.08, 0.16]
    ang = pc.testing.angles_uniform(360)
    sizes, fs, mu_in, mu_out = [], [], [], []
    for n in sizes_sl:
        f = pc.testing.phantom_ellipses(int(n), 10, rng)
        sup = (f > 0).astype(np.float32)
        for sc in scales:
            sizes.append(int(n))
            fs.append((f * rng.uniform(0, 1)).astype(np.float32))
            mu_in.append((0.02 * sup + 0.005).astype(np.float32))
            mu_out.append((sc * sup).astype(np.float32))
    log(f"batched: {len(sizes)} slice-channels, {sum(n * n for n in sizes) / 1e6:.1f} Mpixel")
    otf = args.maps == "otf"
    t0 = time.perf_counter()
    A = pc.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=[np.pi / 2], bp_mode="adjoint",
                                     maps_mode=args.maps)
    torch.cuda.synchronize()
    t_maps = time.perf_counter() - t0
    # the tuned kernels are the persistent batched ones, which the on-the-fly mode does not launch
    tune = None if otf else pc.batch.autotune(A)
    log(f"set-up {t_maps:.1f} s ({A.plan.maps_elems * 4 / 1e9:.1f} GB of maps{' not held' if otf else ''}); "
        f"autotune {tune}")
    clean = A.fp(A.pack_vol(fs))
    gen = torch.Generator(device="cuda").manual_seed(1)
    b = clean + 0.01 * clean.abs().max() * torch.randn(clean.shape, generator=gen, device="cuda")
    sirt = pc.batch.BatchedSIRT()
    pdhg = pc.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2)
    n_sirt, n_pdhg = args.sirt_iters, args.pdhg_iters
    s1 = sirt.prepare(A, b)
    sirt.iterate(A, s1, 11)  # eager step + graph capture
    s2 = pdhg.prepare(A, b)
    pdhg.iterate(A, s2, 11)
    torch.cuda.synchronize()
    log("solver graphs captured")
    t_sirt = pc.batch.time_iterations(lambda k: sirt.iterate(A, s1, k), n_sirt)
    log(f"SIRT-{n_sirt}: {t_sirt:.2f} s")
    t_pdhg = pc.batch.time_iterations(lambda k: pdhg.iterate(A, s2, k), n_pdhg)
    log(f"PDHG-{n_pdhg}: {t_pdhg:.2f} s")
    x = s2["x"]
    y = torch.empty_like(b)
    v = torch.empty_like(x)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    st = torch.cuda.current_stream()
    ev[0].record(st)
    for _ in range(5):
        A.fp(x, y)
    ev[1].record(st)
    for _ in range(5):
        A.bp(y, v)
    ev[2].record(st)
    ev[2].synchronize()
    t_fp = ev[0].elapsed_time(ev[1]) / 5e3
    t_bp = ev[1].elapsed_time(ev[2]) / 5e3
    alg = sum(4.0 * (n * n + 360 * n * n + 360 * n) for n in sizes)
    total = t_sirt + t_pdhg
    samples = 360.0 * sum(n * n for n in sizes)
    return {
        "metric": f"slice-channels/s, batched SIRT-{n_sirt} + PDHG-TV-{n_pdhg} (uneven sizes)",
        "value": len(sizes) / total, "unit": "slice-channels/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.slices} slices x 4 channels, N in {{128..768}}, 360 angles, 1 detector; "
                               f"SIRT-{n_sirt} + PDHG-TV(1e-2, l2)-{n_pdhg}, HIP-graph iterations, "
                               f"{args.maps} maps",
                   "slice_sizes": [int(v) for v in sizes_sl], "maps_gb": A.plan.maps_elems * 4 / 1e9,
                   "maps_mode": args.maps, "s_setup": t_maps, "autotune": tune,
                   "ms_sirt_iteration": 1e3 * t_sirt / n_sirt, "ms_pdhg_iteration": 1e3 * t_pdhg / n_pdhg,
                   "fp_ms": 1e3 * t_fp, "bp_ms": 1e3 * t_bp,
                   # with precomputed maps: algorithmic HBM bytes (slice + maps + sinogram) per second;
                   # on the fly: map samples (pixel-angles) per second, maps built inside the call
                   "fp_gbs": None if otf else alg / t_fp / 1e9, "bp_gbs": None if otf else alg / t_bp / 1e9,
                   "fp_gsamples_per_s": samples / t_fp / 1e9, "bp_gsamples_per_s": samples / t_bp / 1e9},
    }


def run_otf(args) -> dict:
    """On-the-fly maps at the full-scale shape: 1024 x 1024 slices, 1440 angles, one detector; fp and
    adjoint bp with the maps rebuilt per angle block, against the precomputed-maps path on one slice."""
    import torch

    import pycorrectedemissionct_amd as pc

    n, a = 1024, 1440
    rng = np.random.default_rng(20260105)
    ang = pc.testing.angles_uniform(a)
    f = pc.testing.phantom_ellipses(n, 40, rng, n_blobs=8)
    mu_in, mu_out = pc.testing.xrf_attenuation(f, a_in=0.01, a_bg=0.0025, a_out=0.025)
    otf = pc.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, bp_mode="adjoint", maps_mode="otf")
    x = torch.as_tensor(f, device="cuda")
    budgets = {}
    for mib in args.otf_budgets:
        otf._plan().tune("otf_budget_mib", mib)
        otf.fp(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        otf.fp(x)
        torch.cuda.synchronize()
        budgets[mib] = 1e3 * (time.perf_counter() - t0)
        log(f"otf fp with a {mib} MiB scratch budget: {budgets[mib]:.1f} ms")
    best = min(budgets, key=budgets.get)
    otf._plan().tune("otf_budget_mib", best)

    def timed(fn, reps):  # warm call, then the mean of `reps` calls
        out = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t0) / reps

    y, t_fp = timed(lambda: otf.fp(x), 3)
    v, t_bp = timed(lambda: otf.bp(y), 3)
    pre = pc.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, bp_mode="adjoint")
    y2, t_fp_pre = timed(lambda: pre.fp(x), 10)
    _, t_bp_pre = timed(lambda: pre.bp(y2), 10)
    same = bool(torch.equal(y, y2))
    dv = float((pre.bp(y2) - v).abs().max() / v.abs().max())
    # PDHG-KL-TV iterations at this size: precomputed maps, on-the-fly maps built separately for the fp
    # and the bp, and on-the-fly with the fused per-angle-block data step (maps built once)
    b = pc.testing.poisson_data(y2.cpu().numpy(), 1e3, rng)
    nofuse = pc.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, bp_mode="adjoint",
                                        maps_mode="otf")
    nofuse._plan().tune("otf_budget_mib", best)
    nofuse.supports_fused_data_step = lambda: False
    iter_ms, xs_out = {}, {}
    for name, prj in (("precomputed", pre), ("otf_separate", nofuse), ("otf_fused", otf)):
        solver = pc.PDHG(data_term="kl", regularizer=pc.regularizers.Regularizer_TV2D(1e-2))
        st = solver.prepare(prj, b)
        solver.iterate(prj, st, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.iterate(prj, st, args.otf_iters)
        torch.cuda.synchronize()
        iter_ms[name] = 1e3 * (time.perf_counter() - t0) / args.otf_iters
        xs_out[name] = st["x"].clone()
        log(f"PDHG-KL-TV iteration, {name}: {iter_ms[name]:.1f} ms")
    fused_equal = bool(torch.equal(xs_out["otf_fused"], xs_out["otf_separate"]))
    del nofuse
    # the warm start's network at this slice size (8 slices per launch; its cost does not depend
    # on the weights, so they are the seeded initial ones)
    from pycorrectedemissionct_amd import warmstart as ws

    net = ws.UNet(seed=0)
    xs = torch.as_tensor(np.stack([f] * 8), device="cuda")
    ws.warm_start(net, xs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ws.warm_start(net, xs)
    torch.cuda.synchronize()
    t_unet = (time.perf_counter() - t0) / 8
    return {
        "metric": "ms per on-the-fly fp / bp, 1024x1024 slice, 1440 angles",
        "value": 1e3 * (t_fp + t_bp), "unit": "ms per fp+bp pair", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"1024x1024, 1440 angles, 1 detector at 90 deg, maps rebuilt per angle block "
                               f"({best} MiB scratch) instead of 6 GB of precomputed maps",
                   "otf_fp_ms": 1e3 * t_fp, "otf_bp_ms": 1e3 * t_bp, "precomputed_fp_ms": 1e3 * t_fp_pre,
                   "precomputed_bp_ms": 1e3 * t_bp_pre, "otf_over_precomputed_fp": t_fp / t_fp_pre, "otf_fp_ms_by_budget_mib": budgets,
                   "otf_budget_mib": best,
                   # samples the on-the-fly fp computes: the map builder's rotated-frame scan (two beams of
                   # the padded side squared per angle), its composition (two beams per map pixel) and the
                   # projection's own samples (one per pixel and angle)
                   "otf_fp_gsamples_per_s": a * (2 * otf._plan().n_pad ** 2 + 2 * n * n + n * n) / t_fp / 1e9,
                   "otf_fp_equals_precomputed": same, "otf_bp_rel_diff": dv,
                   "pdhg_kl_tv_iteration_ms": iter_ms, "pdhg_iterations_timed": args.otf_iters,
                   "otf_fused_iteration_over_precomputed": iter_ms["otf_fused"] / iter_ms["precomputed"],
                   "otf_separate_iteration_over_precomputed": iter_ms["otf_separate"] / iter_ms["precomputed"],
                   "otf_fused_equals_separate": fused_equal,
                   "precomputed_maps_gb": 4.0 * a * n * n / 1e9, "unet_ms_per_slice": 1e3 * t_unet,
                   "unet_batch": 8},
    }


def run_kernels(args) -> dict:
    """The headline batch's projector kernels alone, for counter runs: the batched fp and exact-adjoint
    bp of --slices 512 x 512 slices at 720 angles and 2 detectors (autotuned shapes) and a 1 GiB device
    copy, each launched --fp-reps times after 3 warm-up launches and timed with HIP events.
    ``--n``, ``--angles`` and ``--detectors 1`` give the other configurations' geometry (one detector at
    90 deg), e.g. the warm-start workload's 256 x 256 x 360."""
    import torch

    import pycorrectedemissionct_amd as pc
    from pycorrectedemissionct_amd import batch as B

    n, n_ang, nb = args.n, args.angles, args.slices
    det = [np.pi / 2, -np.pi / 2][: args.detectors]
    ang = pc.testing.angles_uniform(n_ang)
    fs, mis, mos = [], [], []
    for k in range(nb):
        wl = pc.testing.workload_pdhg_512(seed=20260102 + k, n=n, n_angles=n_ang)
        fs.append(wl["f"]), mis.append(wl["mu_in"]), mos.append(wl["mu_out"])
    A = B.BatchedProjectorXRF([n] * nb, ang, mis, mos, angles_detectors_rad=det, bp_mode="adjoint")
    tune = B.autotune(A)
    x = A.pack_vol(fs)
    y = A.fp(x)
    v = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    def time_launches(fn, reps) -> float:
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t_fp = time_launches(lambda: A.fp(x, y), args.fp_reps)
    t_bp = time_launches(lambda: A.bp(y, v), args.fp_reps)
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    t_copy = time_launches(lambda: dst.copy_(src), args.fp_reps)
    nd = len(det)
    fp_bytes = nb * 4.0 * (n * n + n_ang * nd * n * n + n_ang * nd * n)
    bp_bytes = fp_bytes
    log(f"fp {1e6 * t_fp:.1f} us, bp {1e6 * t_bp:.1f} us, 1 GiB copy {1e6 * t_copy:.1f} us")
    fp_gbs = fp_bytes / t_fp / 1e9
    return {
        "metric": "batched fp / exact-adjoint bp GB/s (algorithmic bytes)", "value": fp_gbs, "unit": "GB/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nb} x {n}x{n} slices, {n_ang} angles, {nd} detector(s), precomputed maps; "
                               f"{args.fp_reps} launches each of the batched fp, the batched adjoint bp and a "
                               f"1 GiB copy", "autotune": tune,
                   "fp_us": 1e6 * t_fp, "bp_us": 1e6 * t_bp, "copy_us": 1e6 * t_copy,
                   "fp_gbs": fp_gbs, "bp_gbs": bp_bytes / t_bp / 1e9, "copy_gbs": 8.0 * src.numel() / t_copy / 1e9,
                   "algorithmic_bytes_per_launch": fp_bytes},
        "roofline": {"bound": "hbm", "achieved": fp_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fp_gbs / HBM_PEAK_GBS, "traffic": None},
    }


def _unet_flops(levels: int, base: int, h: int, w: int) -> float:
    """Multiply-add flops (2 per MAC) of the U-Net's 3x3 convolutions on one h x w image."""
    from pycorrectedemissionct_amd.warmstart import layer_shapes

    total = 0.0
    for name, cin, cout in layer_shapes(levels, base):
        lvl = int(name[-2] if name[-1] in "ab" else name[-1])
        total += 2.0 * 9 * cin * cout * (h >> lvl) * (w >> lvl)
    return total


def run_unet(args) -> dict:
    """The warm-start network (4 levels, 32 base channels, ~2M parameters; bf16 MFMA with float32
    accumulation): inference at 256^2 (batch 32) and 1024^2 (batch 2), and Adam training steps at
    256^2 (batch 8), each timed with HIP events over --steps repetitions; TFLOP/s of the 3x3
    convolutions (2 flops per MAC; a training step counts forward + data gradient + weight
    gradient) against the 2.5 PFLOP/s dense bf16 MFMA peak."""
    import torch

    from pycorrectedemissionct_amd import warmstart

    net = warmstart.UNet(4, 32, seed=0)
    gen = torch.Generator(device="cuda").manual_seed(3)
    st = torch.cuda.current_stream()
    reps = max(1, args.steps // 20)

    def timed(fn, k):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(k):
            fn()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / k

    res = {}
    for n, batch in ((256, 32), (1024, 2)):
        x = torch.rand((batch, n, n), generator=gen, device="cuda")
        t = timed(lambda: net.forward(x), reps)
        fl = batch * _unet_flops(4, 32, n, n)
        res[f"infer_{n}"] = {"batch": batch, "ms": 1e3 * t, "tflops": fl / t / 1e12, "mfma_frac": fl / t / 2.5e15}
        log(f"inference {batch} x {n}^2: {1e3 * t:.2f} ms, {fl / t / 1e12:.1f} TFLOP/s")
    batch, n = 8, 256
    x = torch.rand((batch, n, n), generator=gen, device="cuda")
    tgt = torch.rand((batch, n, n), generator=gen, device="cuda")
    t = timed(lambda: net.train_step(x, tgt), reps)
    # the first convolution (1 input channel) has no data gradient
    fl = batch * (3 * _unet_flops(4, 32, n, n) - 2.0 * 9 * 1 * 32 * n * n)
    res["train_256"] = {"batch": batch, "ms": 1e3 * t, "tflops": fl / t / 1e12, "mfma_frac": fl / t / 2.5e15}
    log(f"training step {batch} x {n}^2: {1e3 * t:.2f} ms, {fl / t / 1e12:.1f} TFLOP/s")
    return {
        "metric": "U-Net TFLOP/s (3x3 convolutions, bf16 MFMA, float32 accumulation)",
        "value": res["infer_256"]["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"U-Net 4 levels x 32 base channels ({net.n_params} parameters); inference "
                               f"32 x 256^2 and 2 x 1024^2, Adam step 8 x 256^2",
                   "reps": reps, **res},
        "roofline": {"bound": "mfma", "achieved": res["infer_256"]["tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": res["infer_256"]["mfma_frac"], "traffic": None},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["headline", "batched", "otf", "unet", "kernels"], default="headline",
                    help="headline: warm-started vs plain batched PDHG at 512^2 (default); batched / otf: the "
                         "uneven batch and the on-the-fly maps; unet: the warm-start network's kernels; kernels: "
                         "the headline batch's fp / bp alone")
    ap.add_argument("--slices", type=int, default=None, help="slices in the batch (headline 2, batched 32)")
    ap.add_argument("--maps", choices=["precomputed", "otf"], default="precomputed",
                    help="batched: attenuation maps held in memory or rebuilt inside every fp / bp")
    ap.add_argument("--sirt-iters", type=int, default=100, help="batched: timed SIRT iterations")
    ap.add_argument("--pdhg-iters", type=int, default=300, help="batched: timed PDHG iterations")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--detectors", type=int, choices=[1, 2], default=2, help="kernels: detectors (1: at 90 deg)")
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--ref-iters", type=int, default=16000)
    ap.add_argument("--tol", type=float, default=5e-2)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--fp-reps", type=int, default=20)
    ap.add_argument("--otf-iters", type=int, default=3, help="otf workload: PDHG iterations timed per variant")
    ap.add_argument("--otf-budgets", type=int, nargs="+", default=[1024, 2048, 4096],
                    help="on-the-fly scratch budgets (MiB) to time; the fastest is kept")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (slices would shard across GPUs with "
              "no exchange); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    torch.cuda.set_device(torch.device("cuda", 0))
    if args.workload == "headline":
        args.slices = args.slices or 2
        res = run_headline(args)
    elif args.workload == "batched":
        args.slices = args.slices or 32
        res = run_batched(args)
    elif args.workload == "unet":
        res = run_unet(args)
    elif args.workload == "kernels":
        args.slices = args.slices or 2
        res = run_kernels(args)
    else:
        res = run_otf(args)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
