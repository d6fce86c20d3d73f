"""Benchmark: PDHG-TV (Poisson data term) reconstructions of 512 x 512 XRF slices.

Workload (one GPU): 512 x 512 float32 phantom of 30 ellipses + 5 Gaussian
blobs, mu_in = 0.02 (f > 0) + 0.005, mu_out = 0.05 (f > 0), 720 angles over
2 pi, two detectors at +-90 deg, Poisson noise at a peak of 1e3 counts;
PDHG with the KL data term and TV lambda = 1e-2.

The projector runs with the exact adjoint as back-projection
(``bp_mode="adjoint"``): with the reference's default back-projection (which
ignores the attenuation maps) PDHG-KL diverges on this workload.

A step is one complete reconstruction of one slice to 1e-3 relative error
against x*, the 2000-iteration plain PDHG solution computed during set-up:
diagonal preconditioner (1 fp + 1 bp) plus the number of PDHG iterations
plain PDHG needs to come within the tolerance of x*. value = slices / s.

Also reported: the fp and the bp (both HBM streams of the precomputed maps)
against the HBM roofline, timed with HIP events on the stream they run on;
the dominant one of the two is the ``roofline`` object. The oracle's C
restatement of the reference CPU path (skimage-semantics fp + bp per
iteration) is timed on a bounded sample on the host cores.
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_err(x, ref) -> float:
    import torch

    return float(torch.linalg.vector_norm((x - ref).double()) / torch.linalg.vector_norm(ref.double()))


def cpu_baseline(prj, k_iters: int, n: int, n_angles: int, budget_s: float = 12.0) -> dict:
    """The oracle's C/OpenMP restatement of the reference CPU path (skimage-semantics fp over the
    precomputed maps + iradon-style bp, all angles and detectors), repeated for about ``budget_s``
    seconds; the time per fp+bp pair is scaled to one solve of ``k_iters`` iterations + preconditioner."""
    from oracle import build_c
    from oracle import skimage_ops as sk

    lib = build_c.load()
    ang = np.asarray(prj.angles_rot_rad, dtype=np.float32)
    degs = sk.rad2deg_f32(ang)
    geom = np.ascontiguousarray(np.stack([sk.radon_matrix(d, n)[[0, 0, 0, 1], [0, 1, 2, 2]] for d in degs]).astype(np.float32))
    th = sk.deg2rad_f32(degs)
    trig = np.ascontiguousarray(np.stack([np.cos(th), np.sin(th)], axis=1).astype(np.float32))
    maps = np.ascontiguousarray(prj.att_vol_angles.cpu().numpy())
    n_det = maps.shape[1]
    w = np.ones((n_angles, n_det), np.float32)
    vol = np.ascontiguousarray(np.random.default_rng(0).random((n, n)).astype(np.float32))
    sino = np.zeros((n_det, n_angles, n), np.float32)
    out = np.zeros((n, n), np.float32)
    t_fp = t_bp = 0.0
    pairs = 0
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        lib.oracle_fp(vol.ctypes.data, maps.ctypes.data, geom.ctypes.data, w.ctypes.data, n, n_angles, n_det,
                      sino.ctypes.data)
        t1 = time.perf_counter()
        lib.oracle_bp_skimage(sino.ctypes.data, trig.ctypes.data, n, n_angles, n_det, out.ctypes.data)
        t2 = time.perf_counter()
        t_fp += t1 - t0
        t_bp += t2 - t1
        pairs += 1
        if t2 - t_start >= budget_s:
            break
    t_iter = (t_fp + t_bp) / pairs
    t_solve = t_iter * (k_iters + 1)
    return {
        "value": 1.0 / t_solve,
        "unit": "slices/s",
        "cores": int(lib.oracle_num_threads()),
        "kind": "port",
        "sample": (f"{pairs} full fp+bp pairs ({n}x{n}, {n_angles} angles x {n_det} detectors, "
                   f"{t_fp + t_bp:.1f} s) of the C/OpenMP oracle: fp {1e3 * t_fp / pairs:.0f} ms, "
                   f"bp {1e3 * t_bp / pairs:.0f} ms per call, scaled to {k_iters} PDHG iterations + "
                   f"preconditioner (elementwise updates not counted)"),
    }


def run_batched(args) -> dict:
    """One GPU's share of the batched workload: 32 slices of sides drawn from {128, 256, 384, 512, 768},
    each with its own 10-ellipse phantom and attenuation and 4 emission channels (mu_out scales 0.02,
    0.04, 0.08, 0.16, channel intensities U[0, 1]), 360 angles, one detector at 90 deg, Gaussian noise
    at 1 % of the max; batched SIRT (100 iterations) then PDHG-TV(1e-2, l2; 300 iterations), HIP-graph
    iterations, autotuned persistent grids."""
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
    t0 = time.perf_counter()
    A = pc.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=[np.pi / 2], bp_mode="adjoint")
    torch.cuda.synchronize()
    t_maps = time.perf_counter() - t0
    tune = pc.batch.autotune(A)
    log(f"maps {t_maps:.1f} s ({A.plan.maps_elems * 4 / 1e9:.1f} GB); autotune {tune}")
    clean = A.fp(A.pack_vol(fs))
    gen = torch.Generator(device="cuda").manual_seed(1)
    b = clean + 0.01 * clean.abs().max() * torch.randn(clean.shape, generator=gen, device="cuda")
    sirt = pc.batch.BatchedSIRT()
    pdhg = pc.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2)
    s1 = sirt.prepare(A, b)
    sirt.iterate(A, s1, 11)  # eager step + graph capture
    s2 = pdhg.prepare(A, b)
    pdhg.iterate(A, s2, 11)
    torch.cuda.synchronize()
    t_sirt = pc.batch.time_iterations(lambda k: sirt.iterate(A, s1, k), 100)
    t_pdhg = pc.batch.time_iterations(lambda k: pdhg.iterate(A, s2, k), 300)
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
    return {
        "metric": "slice-channels/s, batched SIRT-100 + PDHG-TV-300 (uneven sizes)",
        "value": len(sizes) / total, "unit": "slice-channels/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "32 slices x 4 channels, N in {128..768}, 360 angles, 1 detector; "
                               "SIRT-100 + PDHG-TV(1e-2, l2)-300, HIP-graph iterations",
                   "slice_sizes": [int(v) for v in sizes_sl], "maps_gb": A.plan.maps_elems * 4 / 1e9,
                   "s_maps": t_maps, "autotune": tune, "ms_sirt_iteration": 1e3 * t_sirt / 100,
                   "ms_pdhg_iteration": 1e3 * t_pdhg / 300, "fp_ms": 1e3 * t_fp, "bp_ms": 1e3 * t_bp,
                   "fp_gbs": alg / t_fp / 1e9, "bp_gbs": alg / t_bp / 1e9},
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
    y = otf.fp(x)
    otf.bp(y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = otf.fp(x)
    torch.cuda.synchronize()
    t_fp = time.perf_counter() - t0
    t0 = time.perf_counter()
    v = otf.bp(y)
    torch.cuda.synchronize()
    t_bp = time.perf_counter() - t0
    pre = pc.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, bp_mode="adjoint")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y2 = pre.fp(x)
    torch.cuda.synchronize()
    t_fp_pre = time.perf_counter() - t0
    same = bool(torch.equal(y, y2))
    dv = float((pre.bp(y2) - v).abs().max() / v.abs().max())
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
        "config": {"workload": "1024x1024, 1440 angles, 1 detector at 90 deg, maps rebuilt per angle block "
                               "(1 GiB scratch) instead of 6 GB of precomputed maps",
                   "otf_fp_ms": 1e3 * t_fp, "otf_bp_ms": 1e3 * t_bp, "precomputed_fp_ms": 1e3 * t_fp_pre,
                   "otf_fp_equals_precomputed": same, "otf_bp_rel_diff": dv,
                   "precomputed_maps_gb": 4.0 * a * n * n / 1e9, "unet_ms_per_slice": 1e3 * t_unet,
                   "unet_batch": 8},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["pdhg512", "batched", "otf"], default="pdhg512",
                    help="pdhg512: the headline line (default); batched / otf: the batched and on-the-fly paths")
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--ref-iters", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--fp-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (slices would shard across GPUs with "
              "no exchange); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import pycorrectedemissionct_amd as pc

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.workload != "pdhg512":
        res = run_batched(args) if args.workload == "batched" else run_otf(args)
        print(json.dumps(res), flush=True)
        return 0
    n, n_ang = args.n, args.angles
    log(f"set-up: {n}x{n}, {n_ang} angles, 2 detectors")
    wl = pc.testing.workload_pdhg_512(n=n, n_angles=n_ang)
    prj = pc.ProjectorAttenuationXRF([n, n], wl["angles"], att_in=wl["mu_in"], att_out=wl["mu_out"],
                                     angles_detectors_rad=wl["det"], bp_mode="adjoint")
    f = torch.as_tensor(wl["f"], device=dev)
    sino = prj.fp(f)
    b_np = pc.testing.poisson_data(sino.cpu().numpy(), wl["peak"], wl["rng"])
    b = torch.as_tensor(b_np, device=dev)
    reg = pc.regularizers.Regularizer_TV2D(1e-2)
    solver = pc.PDHG(data_term="kl", regularizer=reg)

    # converged reference and iterations-to-tolerance of plain PDHG (untimed set-up)
    s = solver.prepare(prj, b)
    log(f"reference: {args.ref_iters} plain PDHG iterations")
    solver.iterate(prj, s, args.ref_iters)
    x_star = s["x"].clone()
    s = solver.prepare(prj, b)
    k_tol = None
    done = 0
    while done < args.ref_iters:
        solver.iterate(prj, s, 10)
        done += 10
        if rel_err(s["x"], x_star) <= args.tol:
            k_tol = done
            break
    if k_tol is None:
        k_tol = args.ref_iters
    log(f"plain PDHG reaches {args.tol:g} relative error in {k_tol} iterations")

    def one_solve():
        st = solver.prepare(prj, b)
        solver.iterate(prj, st, k_tol)
        return st

    for _ in range(args.warmup):
        one_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = one_solve()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    err = rel_err(st["x"], x_star)
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    log(f"timed {args.steps} solves: {ms_per_step:.1f} ms/slice, final rel err {err:.2e}")

    # fp and bp kernels over the precomputed maps, HIP events on the stream they are launched on
    stream = torch.cuda.current_stream()
    n_det = 2
    out = torch.empty((n_det, n_ang, n), dtype=torch.float32, device=dev)
    vol_t = torch.empty_like(f)
    plan = prj._plan()
    maps = prj.att_vol_angles
    xr = s["x"]
    bp_mode = pc._native.CHIP_BP_ADJOINT

    def time_launches(fn) -> float:
        for _ in range(3):
            fn()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.fp_reps):
            fn()
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1) / args.fp_reps * 1e-3

    t_fp = time_launches(lambda: plan.fp(xr, maps, out, 1))
    t_bp = time_launches(lambda: plan.bp(out, maps, vol_t, 1, bp_mode))
    # algorithmic HBM bytes per launch: fp reads x and the maps and writes the sinogram;
    # the adjoint bp reads the sinogram and the maps and writes x
    fp_bytes = 4.0 * (n * n + n_ang * n_det * n * n + n_ang * n_det * n)
    bp_bytes = 4.0 * (n_ang * n_det * n + n_ang * n_det * n * n + n * n)
    fp_gbs, bp_gbs = fp_bytes / t_fp / 1e9, bp_bytes / t_bp / 1e9
    log(f"fp {t_fp * 1e6:.1f} us ({fp_gbs:.0f} GB/s), bp {t_bp * 1e6:.1f} us ({bp_gbs:.0f} GB/s)")

    # the configuration's second solve: MLEM, 300 iterations, through the solver API
    mlem = pc.MLEM()
    mlem(prj, b, iterations=5)
    torch.cuda.synchronize()
    tm0 = time.perf_counter()
    mlem(prj, b, iterations=300)
    torch.cuda.synchronize()
    t_mlem = time.perf_counter() - tm0
    log(f"MLEM 300 iterations: {t_mlem * 1e3:.1f} ms")

    # device-to-device copy of 1 GiB (read 1 GiB, write 1 GiB): the bandwidth a plain stream reaches here
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    t_copy = time_launches(lambda: dst.copy_(src))
    copy_gbs = 2.0 * 4 * src.numel() / t_copy / 1e9
    del src, dst
    log(f"copy 1 GiB: {t_copy * 1e6:.1f} us ({copy_gbs:.0f} GB/s)")

    kernels = {
        "fp": dict(kernel="fp_stream_kernel<FpTile<32, 32, 48, 13, 52, 1>, 2, true, false>", t=t_fp, bytes=fp_bytes, gbs=fp_gbs,
                   traffic_file="fp_stream_512x720x2_traffic.json"),
        "bp": dict(kernel="bp_adjoint_tile_kernel<2, true>", t=t_bp, bytes=bp_bytes, gbs=bp_gbs,
                   traffic_file="bp_adjoint_512x720x2_traffic.json"),
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

    result = {
        "metric": "slices/s to 1e-3 rel err, warm-start vs plain PDHG-TV; fp+bp HBM GB/s",
        "value": 1e3 / ms_per_step,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "PDHG-KL-TV(1e-2) 512x512 XRF slice, 720 angles over 2pi, 2 detectors at +-90 deg, "
                        "Poisson peak 1e3, exact-adjoint bp; to 1e-3 rel err vs the 2000-iteration solution; "
                        "plain PDHG (no warm start)",
            "n": n, "angles": n_ang, "detectors": n_det, "pdhg_iterations": k_tol, "final_rel_err": err,
            "ms_per_iteration": ms_per_step / (k_tol + 1), "mlem300_ms": 1e3 * t_mlem,
            "fp_us": t_fp * 1e6, "bp_us": t_bp * 1e6,
            "fp_plus_bp_gbs": (fp_bytes + bp_bytes) / (t_fp + t_bp) / 1e9,
        },
        "roofline": {
            "bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": dom["kernel"], "algorithmic_bytes_per_launch": dom["bytes"],
            "copy_gbs_measured": copy_gbs, "frac_of_copy": dom["gbs"] / copy_gbs,
            "other": {k: {"kernel": v["kernel"], "achieved": v["gbs"], "frac": v["gbs"] / HBM_PEAK_GBS,
                          "us": v["t"] * 1e6} for k, v in kernels.items() if v is not dom},
        },
    }
    if not args.no_cpu:
        log(f"cpu baseline (oracle C restatement, about {args.cpu_seconds:g} s)")
        result["cpu_baseline"] = cpu_baseline(prj, k_tol, n, n_ang, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
