"""Benchmark of the plain method on one MI355X.

Workload (the full-size single-GPU configuration): 600 off-axis holograms
1024x1024 uint16 of a seeded 5-12 ellipsoid phantom, Fermat-spiral
illumination filling 0.95 NA, Poisson shot noise at 20k photons peak,
sideband crop 256x256, the count-normalised 512^3 complex64 spectrum (its
hit band; counts are not stored), 3-D inverse FFT to float32 index and
absorption volumes. One step = holograms already in HBM -> index volume
(mtd_reconstruct_plain: both programs' compute).

Prints one JSON line. value = holograms/s of the end-to-end step
(A / step time); ms_per_step = ms per 512^3 volume.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("holograms/s and ms per 512³ volume; plain vs adjoint-TV at equal "
          "index RMSE")
PEAK_HBM_GBS = 8000.0

# cfg: N, NXMAX, V, A
WORKLOADS = {
    "cfg2": dict(n=1024, nxmax=128, v=512, A=600, seed=2024),
    # differentiation form: plain vs adjoint-TV index RMSE against the phantom
    "diff": dict(n=768, nxmax=64, v=256, A=400, seed=2025),
    # large combined configuration (plain method only): --workload large
    "large": dict(n=2048, nxmax=256, v=1024, A=1000, seed=2026),
}
# adjoint-TV settings of the quality comparison (objective in index units:
# 1/2 ||A dn - y||_W^2 / ||A^H W A|| + lam * TV_eps(dn)); the RMSE curve is
# followed to TV_CURVE iterations in an untimed rerun
TV_ITERS, TV_LAM, TV_EPS, TV_CURVE = 20, 1e-2, 2e-2, 50
LAM, N0, NA = 6.33e-7, 1.333, 1.2


def geometry(w):
    n, nx = w["n"], w["nxmax"]
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    return n, nx, w["v"], w["A"], pix, cx, cy


def round_away(x):
    import numpy as np
    t = np.trunc(x)
    return t + np.where(np.abs(x - t) >= 0.5, np.sign(x), 0.0)


def hit_band(pos, nx, V, R, coef_na):
    """Height band [lo, hi] that holds every hit of the step (the library
    computes the same band on the device): per non-skipped angle, heights
    between the pupil rim and centre, h(r2) = sqrt(R^2 - r2)."""
    import numpy as np
    A = pos.size // 2
    fx = pos[:A].astype(np.int64) - nx
    fy = pos[A:].astype(np.int64) - nx
    R2 = R * R
    r2tab = 2 * nx * nx + 1
    r2max = max(r for r in range(r2tab) if r < coef_na * nx * nx)
    skip = round_away((R2 - fx * fx) - fy * fy) < 0
    r2 = fx * fx + fy * fy
    fz = np.where(~skip & (r2 < r2tab), np.sqrt(np.maximum(R2 - r2, 0.0)), 0.0)
    act = ~skip
    if not act.any():
        return 1, 0
    lo = int(round_away(np.sqrt(R2 - r2max) - fz[act]).min()) + V // 2
    hi = int(round_away(R - fz[act]).max()) + V // 2
    return max(lo, 0), min(hi, V - 1)


def kernel_bytes(name, n, C, V, A, P, M, B, b_in=2):
    """Algorithmic (compulsory) bytes of one step per kernel, from the
    per-hologram stage bytes of the measurement definition: S1 = b_in N^2 +
    8 C^2 (the hologram read by the row pass, the crop written by the column
    pass; the N x C row intermediate between them is not compulsory), S2 =
    16 C^2 (the crop read by the demodulation's row pass -- in the fused
    field stage the crop's column transform Q, the same size, read by the
    one row kernel -- and the spectrum written by the forward column pass;
    the passes in between move intermediates), S3 = 8 P (pupil samples read by the gather). Per volume
    the fused plain method writes only the B planes of the hit band of the
    count-normalised spectrum (8 B V^2; no counts), and the finalisation
    reads that band and writes the index and absorption volumes (8 V^3).
    Kernels that only move intermediates count 0 (their rate is reported from
    the counters instead)."""
    table = {
        "holo_rows": A * b_in * n * n,
        "holo_cols": A * 8 * C * C,
        "demod_rows": A * 8 * C * C,
        "fwd_rows_field": A * 8 * C * C,
        "fwd_cols": A * 8 * C * C,
        "ewald_gather": A * 8 * P + 8 * B * V * V,
        "vol_x": 8 * B * V * V,
        "vol_z": 8 * V ** 3,
    }
    return table.get(name)


def stage_bytes(n, C, V, A, P, B, b_in=2):
    """Per stage: the measurement definition's bytes (S1, S2, S3 per
    hologram; accumulation 12 V^3 of spectrum + counts; S4 = 16 V^3) and the
    bytes this design must move (banded spectrum, no counts, index and
    absorption written)."""
    s1, s2, s3 = A * (b_in * n * n + 8 * C * C), A * 16 * C * C, A * 8 * P
    return {
        "S1": {"definition": s1, "design": s1},
        "S2": {"definition": s2, "design": s2},
        "S3_acc": {"definition": s3 + 12 * V ** 3, "design": s3 + 8 * B * V * V},
        "S4": {"definition": 16 * V ** 3, "design": 8 * B * V * V + 8 * V ** 3},
    }


def adjoint_quality(ctx, torch, np, m, synth, wname="diff", n_iter=TV_ITERS,
                    curve_iters=TV_CURVE, graph=False):
    """The differentiation form against the plain method on its workload
    (default: 400 off-axis holograms 768^2 uint16 of a seeded ellipsoid
    phantom, Fermat spiral at 0.95 NA, Poisson noise, crop 128^2, 256^3 index
    volume; "large": 1000 x 2048^2 -> crop 512^2 -> 1024^3), the shipped
    recon.txt field settings (BORN 0 Rytov, DEROUL 1 Volkov unwrapping,
    C_ABER 1). Plain = mtd_reconstruct_plain; adjoint-TV = the same
    preprocessing, per-angle spectra y, then n_iter FISTA iterations on
    1/2 ||A dn - y||_W^2 / kappa + lam TV(dn) with dn >= 0 from the plain
    start, run by the library (mtd_tv_solve); graph: the iterations
    captured as one HIP graph (its capture and instantiation, a few ms, are
    inside the timed call). Index RMSE against the voxelized phantom; GPU
    times of both."""
    from mtd_transmission_amd import adjoint as adj
    w = WORKLOADS[wname]
    n, nx, V, A, pix, cx, cy = geometry(w)
    C = 2 * nx
    rng = np.random.default_rng(w["seed"])
    ph = synth.make_phantom(w["seed"], int(rng.integers(5, 13)))
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph,
                                       seed=w["seed"] + 1)
    g = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix, 1.0, 1.0, 1.0, 1.0, 0, 1, 1)
    re = torch.empty((V, V, V), dtype=torch.float32, device="cuda")
    im = torch.empty_like(re)
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    fields = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(fields)
    W = torch.empty((A, C, C), dtype=torch.float32, device="cuda")
    vol = torch.empty((V, V, V), dtype=torch.complex64, device="cuda")
    cnt = torch.empty((V, V, V), dtype=torch.int32, device="cuda")

    def plain():
        ctx.reconstruct_plain(g, holos, mask, re, im, pos)

    def tv_setup():
        ctx.preprocess(g, holos, mask, pos, fields)
        ctx.field_to_spectrum(g, fields, pos, y)
        ctx.cap_weights(g, pos, W)
        ctx.ewald_scatter(g, y, pos, vol, cnt)
        ctx.finalize(g, vol, cnt, re, im, None)

    def tv():
        tv_setup()
        kappa = (1.0 / (V * g.delta_f)) ** 6 * V ** 3  # ||A^H W A||, closed form
        return adj.reconstruct_tv(ctx, g, pos, y, n_iter=n_iter, lam=TV_LAM, eps=TV_EPS,
                                  dn0=re, kappa=kappa, weights=W, fused=True, graph=graph)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps, r

    ms_plain, _ = timed(plain, 5)
    truth = synth.voxelize_torch(ph, V, 1.0 / (V * g.delta_f))
    rmse = lambda x: float((x.double() - truth.double()).pow(2).mean().sqrt())
    # the phantom's support dilated by 2 voxels (a 5^3 max filter)
    sup = torch.nn.functional.max_pool3d((truth > 0).float()[None, None], 5, 1, 2)[0, 0] > 0
    rmse_sup = lambda x: float((x.double() - truth.double())[sup].pow(2).mean().sqrt())
    r_plain = rmse(re)
    rs_plain = rmse_sup(re)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    alloc0 = torch.cuda.memory_allocated()
    ms_tv, res = timed(tv, 2)
    torch.cuda.synchronize()
    free_b, total_b = torch.cuda.mem_get_info()
    # the solve's device memory: torch tensors (inputs, spectra, weights,
    # iterates) at their peak during the timed solves, and the library's own
    # workspaces (device memory in use that torch does not hold)
    mem = {"torch_allocated_before_solve": round(alloc0 / 1e9, 2),
           "torch_peak_during_solve": round(torch.cuda.max_memory_allocated() / 1e9, 2),
           "library_workspaces": round((total_b - free_b - torch.cuda.memory_reserved()) / 1e9,
                                       2),
           "device_total": round(total_b / 1e9, 1)}
    r_tv = rmse(res.dn)
    rs_tv = rmse_sup(res.dn)
    # the RMSE along the iterations (an untimed rerun with a monitor), and the
    # iterations that fit in 10x the plain method's time
    curve = {}
    kappa = (1.0 / (V * g.delta_f)) ** 6 * V ** 3
    adj.reconstruct_tv(ctx, g, pos, y, n_iter=curve_iters, lam=TV_LAM, eps=TV_EPS, dn0=re,
                       kappa=kappa, weights=W,
                       monitor=lambda k, x: curve.__setitem__(k, rmse(x)))
    # the configuration's literal settings: lam_TV in {1e-4, 1e-3}, 100
    # iterations (lam in index units of the kappa-normalised objective)
    sweep = {}
    if wname == "diff":
        for lam in (1e-4, 1e-3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = adj.reconstruct_tv(ctx, g, pos, y, n_iter=100, lam=lam, eps=TV_EPS, dn0=re,
                                   kappa=kappa, weights=W, fused=True)
            torch.cuda.synchronize()
            sweep[f"{lam:g}"] = {"iterations": 100, "index_rmse": round(rmse(r.dn), 6),
                                 "rmse_reduction": round(1 - rmse(r.dn) / r_plain, 4),
                                 "ms_solver": round((time.perf_counter() - t0) * 1e3, 2)}
    # The configuration's objective ||A(n) - y||^2 + lam_cfg TV(n), read with A
    # normalised to ||A^H A|| = 1 (in SI units ||A^H A|| is about 1e-40, so a
    # fixed numeric lam is only meaningful for a unit-free operator): with
    # A^ = A / sqrt(kappa), y^ = y / sqrt(kappa), ||A^ n - y^||^2 + lam_cfg TV
    # = 2 (1/2 ||A n - y||^2 / kappa + (lam_cfg / 2) TV), the solver's
    # objective with W = 1 at lam = lam_cfg / 2. Run at the configuration's two
    # values, 100 iterations each, with the RMSE along the way; the headline
    # (count-weighted) lam = 1e-2 is lam_cfg = 2e-2 of the weighted objective.
    cfg_obj = {}
    if wname == "diff":
        kappa1 = adj.normal_norm(ctx, g, pos, None)
        cfg_obj = {"mapping": "lam_solver = lam_cfg / 2 (objective normalised by "
                              "||A^H A||; unweighted, W = 1)",
                   "kappa_unweighted": kappa1, "kappa_weighted": kappa,
                   "headline_lam_cfg_weighted": 2 * TV_LAM, "runs": {}}
        for lam_cfg in (1e-4, 1e-3):
            cur = {}
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = adj.reconstruct_tv(ctx, g, pos, y, n_iter=100, lam=lam_cfg / 2, eps=TV_EPS,
                                   dn0=re, kappa=kappa1, weights=None,
                                   monitor=lambda k, x: cur.__setitem__(k, rmse(x)))
            torch.cuda.synchronize()
            cfg_obj["runs"][f"{lam_cfg:g}"] = {
                "lam_solver": lam_cfg / 2, "iterations": 100,
                "index_rmse": round(rmse(r.dn), 6),
                "rmse_reduction": round(1 - rmse(r.dn) / r_plain, 4),
                "best_rmse_reduction": round(1 - min(cur.values()) / r_plain, 4),
                "rmse_by_iteration": {str(k): round(cur[k], 6) for k in (5, 10, 20, 50, 100)
                                      if k in cur},
                "ms_with_monitor": round((time.perf_counter() - t0) * 1e3, 1)}
    setup_ms, _ = timed(lambda: tv_setup(), 3)
    per_iter = max((ms_tv - setup_ms) / n_iter, 1e-6)
    k10 = int(max(0, min(curve_iters, (10 * ms_plain - setup_ms) // per_iter)))
    r10 = curve.get(k10, r_plain)
    print(f"[bench] adjoint-TV: index RMSE plain {r_plain:.5f} ({ms_plain:.2f} ms), "
          f"adjoint-TV {r_tv:.5f} ({ms_tv:.1f} ms); {k10} iterations in 10x plain time: "
          f"{r10:.5f}", file=sys.stderr, flush=True)
    return {"workload": f"{A} x {n}^2 uint16 holograms -> crop {C}^2 -> {V}^3 index; "
                        "BORN 0, DEROUL 1 (Volkov), C_ABER 1",
            "index_rmse_plain": round(r_plain, 6), "index_rmse_adjoint_tv": round(r_tv, 6),
            "rmse_reduction": round(1 - r_tv / r_plain, 4),
            "support_dilated_2": {"voxels": int(sup.sum()),
                                  "index_rmse_plain": round(rs_plain, 6),
                                  "index_rmse_adjoint_tv": round(rs_tv, 6),
                                  "rmse_reduction": round(1 - rs_tv / rs_plain, 4)},
            "ms_plain": round(ms_plain, 3), "ms_adjoint_tv": round(ms_tv, 2),
            "time_ratio": round(ms_tv / ms_plain, 1),
            "iterations": n_iter, "ms_per_iteration": round(per_iter, 3),
            "solver": "mtd_tv_solve" + (", all iterations captured as one HIP graph"
                                        if graph else ", eager launches"),
            "rmse_by_iteration": {str(k): round(curve[k], 6) for k in (5, 10, 15, 20, 30, 50)
                                  if k in curve},
            "iterations_within_10x_plain": k10,
            "rmse_reduction_within_10x_plain": round(1 - r10 / r_plain, 4),
            "lam_tv": TV_LAM, "eps_tv": TV_EPS,
            "lam_sweep_100_iterations": sweep,
            "config_objective": cfg_obj,
            "device_memory_GB": mem,
            "weights": "1/(q^2 hits)", "start": "plain"}


def cpu_baseline_child(path_holos, path_out, threads):
    """Runs in a child process: the reference CPU path's FP64 arithmetic at
    its stage boundaries over every hologram (oracle/cpu_pipeline.py), the
    FFTs on `threads` threads, the host loops serial."""
    import numpy as np
    from oracle import cpu_pipeline, ref
    d = np.load(path_holos)
    holos, mask = d["holos"], d["mask"]
    n, nx, v, A, pix, cx, cy = geometry(WORKLOADS["cfg2"])
    g = ref.derive_geom(n, nx, v, cx, cy, NA, N0, LAM, pix, 1, 1, 1)
    t0 = time.perf_counter()
    _, _, t, name = cpu_pipeline.run(holos, g, mask, threads)
    t["total"] = time.perf_counter() - t0
    json.dump({"A": int(holos.shape[0]), "stages_s": t, "fft": name}, open(path_out, "w"))


def host_cpu():
    """nproc and the CPU model of this host (lscpu), and the threads the
    baseline may use: at most 16, the host cores that go with one GPU."""
    import shutil
    nproc = os.cpu_count() or 1
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    model = ""
    if shutil.which("lscpu"):
        r = subprocess.run(["lscpu"], capture_output=True, text=True)
        for line in r.stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    return nproc, model


def cpu_baseline(holos_dev, mask, A):
    """The reference CPU path on this host over all A holograms, at the
    reference's setting (NB_THREADS 4 for the FFTs) and with every core of
    the host's share (min(nproc, 16))."""
    import numpy as np
    import tempfile
    nproc, model = host_cpu()
    tmp = tempfile.mkdtemp(prefix="mtd_cpu_")
    ph = os.path.join(tmp, "in.npz")
    np.savez(ph, holos=holos_dev[:A].cpu().numpy(), mask=mask)
    runs = {}
    for label, th in (("NB_THREADS_4", 4), ("nproc", min(nproc, 16))):
        po = os.path.join(tmp, f"out_{th}.json")
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1",
                   MKL_NUM_THREADS="1", HIP_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
        print(f"[bench] CPU baseline: {A} holograms, FFT threads {th}", file=sys.stderr,
              flush=True)
        subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", ph, po,
                        str(th)], check=True, env=env)
        r = json.load(open(po))
        os.remove(po)
        st = r["stages_s"]
        compute = st["T_pre1"] + st["T_pre2"] + st["T_rec1"] + st["T_rec2"]
        runs[label] = {"threads": th, "fft": r["fft"],
                       "stages_s": {k: round(v, 3) for k, v in st.items()},
                       "ms_per_volume": round(compute * 1e3, 1),
                       "holograms_per_s": round(r["A"] / compute, 3),
                       "holograms_per_s_s1_to_s3": round(
                           r["A"] / (st["T_pre1"] + st["T_pre2"] + st["T_rec1"]), 3)}
    os.remove(ph)
    os.rmdir(tmp)
    return runs, nproc, model


def copy_bandwidth(torch, nbytes=1 << 30, reps=10):
    """Device-to-device copy of nbytes (read + write): the achievable HBM
    figure on this GPU, and the reference kernel for counter scaling."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del src, dst
    return {"bytes_per_copy": 2 * nbytes, "GBps": round(2 * nbytes / dt / 1e9, 1)}


def batch_shapes(seed=2027, count=32):
    """The batched workload: per acquisition an angle count drawn from
    50-800, a hologram side from {512, 768, 1024} and a crop side from
    128-384 in steps of 32, V = 2C; crops are drawn among those whose window
    (carrier +- NXMAX) stays inside the hologram, NXMAX <= 0.2 N."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        A = int(rng.integers(50, 801))
        n = int(rng.choice([512, 768, 1024]))
        crops = [c for c in range(128, 385, 32) if c // 2 <= 0.2 * n]
        C = int(rng.choice(crops))
        out.append(dict(n=n, nxmax=C // 2, v=2 * C, A=A))
    return out


def run_batched(args):
    """Batched form: 32 acquisitions of uneven shapes through
    mtd_reconstruct_batch, eagerly and as one replayed HIP graph (the timed
    steps); per-acquisition latency p50/p99 from the library's wall-clock
    stamps over the timed replays."""
    import numpy as np
    import torch

    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth

    shapes = batch_shapes()
    geoms, holos, masks = [], [], []
    t0 = time.perf_counter()
    for i, s in enumerate(shapes):
        w = dict(s, seed=3000 + i)
        n, nx, V, A, pix, cx, cy = geometry(w)
        ph = synth.make_phantom(w["seed"], 2)
        kx, ky = synth.scan_fermat(A, nx, 0.95)
        h, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph,
                                       seed=w["seed"] + 1, lowpass=False)
        geoms.append(m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix))
        holos.append(h)
        masks.append(mask)
    torch.cuda.synchronize()
    total_A = sum(s["A"] for s in shapes)
    print(f"[bench] synthesized {len(shapes)} acquisitions, {total_A} holograms in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    re = [torch.empty((g.dim_final,) * 3, device="cuda") for g in geoms]
    im = [torch.empty_like(r) for r in re]
    pos = [torch.empty((2 * h.shape[0],), dtype=torch.int32, device="cuda") for h in holos]
    import tempfile
    cache = args.tune_cache or os.path.join(tempfile.mkdtemp(prefix="mtd_tune_"), "tune.txt")
    ctx = m.Context(0)
    ctx.set_tune_cache(cache)
    # eager batch (no graph) for comparison, after a warm-up (which tunes
    # every shape not in the cache file)
    t0 = time.perf_counter()
    ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=False, autotune=True)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=False, autotune=True)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # capture, then timed replays
    ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, autotune=True)
    for _ in range(args.warmup):
        ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, autotune=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True, autotune=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    print(f"[bench] {args.steps} graph replays, {ms:.3f} ms per batch (eager {eager_ms:.3f})",
          file=sys.stderr, flush=True)
    lat = []
    for _ in range(args.steps):  # per-acquisition stamps (each call waits)
        lat.extend(ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=True,
                                         autotune=True, times=True).tolist())
    lat = np.asarray(lat)
    per_acq = [dict(s, ms=round(float(t), 4)) for s, t in zip(shapes, lat[-len(shapes):])]
    scheds = {}
    for g in geoms:
        sc = ctx.batch_tuned_schedule(g.n, g.C, g.dim_final)
        scheds[f"{g.n}x{g.C}x{g.dim_final}"] = (
            [sc[k] for k in ("holo_chunk", "gather_window", "gather_batch", "vol_rows_tile",
                             "vol_cols_tile", "from_cache")] if sc else None)
    ctx.close()
    # a second context on the same cache file: the schedules are read back,
    # nothing is re-tuned
    ctx = m.Context(0)
    ctx.set_tune_cache(cache)
    cached = sum(1 for g in geoms if (ctx.batch_tuned_schedule(g.n, g.C, g.dim_final) or {})
                 .get("from_cache") == 1)
    t0 = time.perf_counter()
    ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=False, autotune=True)
    torch.cuda.synchronize()
    second_first_ms = (time.perf_counter() - t0) * 1e3
    ctx.close()
    small = batched_small_graph_gain(args, m, synth, torch, np)
    out = {
        "metric": METRIC, "value": round(total_A / (ms / 1e3), 2), "unit": "holograms/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched form: {len(shapes)} acquisitions, angles 50-800, "
                               "holograms {512,768,1024}^2 uint16, crops 128-384 (V = 2C), "
                               "one captured HIP graph per step",
                   "acquisitions": len(shapes), "holograms": total_A},
        "batched": {"p50_ms": round(float(np.percentile(lat, 50)), 4),
                    "p99_ms": round(float(np.percentile(lat, 99)), 4),
                    "eager_ms_per_batch": round(eager_ms, 4),
                    "graph_ms_per_batch": round(ms, 4),
                    "tuned_schedules": {"key": "hologram side x crop x volume side",
                                        "value": "[holograms per hologram launch pair, "
                                                 "gather window height, gather load batch, "
                                                 "volume row tile variant, volume column "
                                                 "tile variant, read from cache]", **scheds},
                    "tune_cache": {"file": cache, "first_call_s_with_tuning": round(tune_s, 2),
                                   "second_context_shapes_from_cache": cached,
                                   "second_context_first_call_ms": round(second_first_ms, 2)},
                    "graph_gain_small_shapes": small,
                    "acquisitions": per_acq},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def batched_small_graph_gain(args, m, synth, torch, np, count=32):
    """Where a captured graph could pay: many small acquisitions (256^2
    holograms, crop 64, 128^3, 24 angles), whose kernels last a few
    microseconds each, eager launches against one graph replay: device time
    per batch and the host time until the call returns."""
    geoms, holos, masks = [], [], []
    for i in range(count):
        w = dict(n=256, nxmax=32, v=128, A=24, seed=5000 + i)
        n, nx, V, A, pix, cx, cy = geometry(w)
        kx, ky = synth.scan_fermat(A, nx, 0.95)
        # the phantom scaled into the small field of view (0.45 n pix)
        ph = synth.scale_phantom(synth.make_phantom(w["seed"], 2), 0.45 * n * pix / 33.8e-6)
        h, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph,
                                       seed=w["seed"] + 1, lowpass=False)
        geoms.append(m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix))
        holos.append(h)
        masks.append(mask)
    re = [torch.empty((g.dim_final,) * 3, device="cuda") for g in geoms]
    im = [torch.empty_like(r) for r in re]
    pos = [torch.empty((2 * h.shape[0],), dtype=torch.int32, device="cuda") for h in holos]
    ctx = m.Context(0)
    res = {}
    for graph in (False, True):
        for _ in range(2):
            ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        host = 0.0
        for _ in range(args.steps):
            h0 = time.perf_counter()
            ctx.reconstruct_batch(geoms, holos, masks, re, im, pos, graph=graph)
            host += time.perf_counter() - h0  # the call returns before the GPU is done
        torch.cuda.synchronize()
        tag = "graph" if graph else "eager"
        res[f"{tag}_ms"] = round((time.perf_counter() - t0) * 1e3 / args.steps, 4)
        res[f"{tag}_host_ms_per_call"] = round(host * 1e3 / args.steps, 4)
    ctx.close()
    res["acquisitions"] = count
    res["shape"] = "256^2 holograms, crop 64, 128^3, 24 angles"
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the PCIe-inclusive leg (host stacks upload in smaller "
                         "chunks; profiles averaging per-dispatch counters skip it)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline (the reference CPU path over every hologram)")
    ap.add_argument("--cpu-child", nargs=3, help=argparse.SUPPRESS)
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the plain vs adjoint-TV comparison (differentiation form)")
    ap.add_argument("--tune-cache", default=None,
                    help="batched workload: the autotuner's cache file (default: a new "
                         "temporary file, so the first context tunes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device copies of 1 GiB (reference kernel for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scaling) and exit")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_baseline_child(args.cpu_child[0], args.cpu_child[1], int(args.cpu_child[2]))
        return
    if args.copy_probe:
        import torch
        print(json.dumps(copy_bandwidth(torch, reps=3)))
        return
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(it is measured on a single MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        run_batched(args)
        return

    import numpy as np
    import torch

    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth

    w = WORKLOADS[args.workload]
    n, nx, V, A, pix, cx, cy = geometry(w)
    C = 2 * nx
    rng = np.random.default_rng(w["seed"])
    n_ell = int(rng.integers(5, 13))
    phantom = synth.make_phantom(w["seed"], n_ell)
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    t0 = time.perf_counter()
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, phantom,
                                       seed=w["seed"] + 1)
    torch.cuda.synchronize()
    print(f"[bench] synthesized {A} holograms {n}^2 uint16 in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)

    g = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix)
    ctx = m.Context(0)
    re = torch.empty((V, V, V), dtype=torch.float32, device="cuda")
    im = torch.empty_like(re)
    pos = torch.empty((2 * A,), dtype=torch.int32, device="cuda")
    for _ in range(args.warmup):
        ctx.reconstruct_plain(g, holos, mask, re, im, pos)
    torch.cuda.synchronize()

    # per-kernel breakdown, in untimed steps: every kernel between a pair of
    # HIP events on the library's stream (each pair delays the next dispatch
    # by a few microseconds, so the timed region below carries only the
    # dominant kernel's)
    P = int(sum(1 for y in range(-nx, nx) for x in range(-nx, nx) if x * x + y * y < nx * nx))
    M = int(((mask[::5, ::5]) > 45).sum())
    lo, hi = hit_band(pos.cpu().numpy(), nx, V, g.rayon, g.coef_na)
    B = max(hi - lo + 1, 0)
    ctx.timing_reset()
    ctx.set_timing(True)
    for _ in range(args.steps):
        ctx.reconstruct_plain(g, holos, mask, re, im, pos)
    torch.cuda.synchronize()
    ctx.set_timing(False)
    kt_all = ctx.timing()
    counters = {}
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and args.workload == "cfg2":  # measured on that workload
        counters = json.load(open(tpath))
    per_kernel = {}
    for name, (launches, tot) in kt_all.items():
        b = kernel_bytes(name, n, C, V, A, P, M, B)
        lps = launches / args.steps
        avg = tot / launches
        per_kernel[name] = {"ms_per_launch": round(avg, 5), "launches_per_step": lps,
                            "ms_per_step": round(tot / args.steps, 4)}
        if b:
            per_kernel[name]["GBps_algorithmic"] = round(b / lps / (avg * 1e-3) / 1e9, 1)
        if isinstance(counters.get(name), (int, float)):
            per_kernel[name]["GBps_counters"] = round(counters[name] / (avg * 1e-3) / 1e9, 1)
    # stage times from the breakdown: S1 hologram stage, S2 fields and
    # per-angle spectra, S3 + accumulation (Ewald gather and its geometry),
    # S4 finalisation (3-D inverse FFT to the index volume)
    stage_of = {"holo_rows": "S1", "holo_cols": "S1", "argmax": "S2", "argmax_fin": "S2",
                "demod_rows": "S2", "demod_cols": "S2", "lsq": "S2", "fwd_rows_field": "S2",
                "fwd_cols": "S2", "sample_rows": "S2",
                "angle_geo": "S3_acc", "ewald_band": "S3_acc", "ewald_tiles": "S3_acc",
                "ewald_gather": "S3_acc", "vol_x": "S4", "vol_y": "S4", "vol_z": "S4"}
    stages = {}
    for name, (launches, tot) in kt_all.items():
        st = stage_of.get(name, "other")
        stages[st] = round(stages.get(st, 0.0) + tot / args.steps, 4)
    # each stage's bytes over its kernels' time, against the 8 TB/s peak:
    # the measurement definition's bytes and this design's (banded spectrum)
    sb = stage_bytes(n, C, V, A, P, B)
    stage_roofline = {}
    for st, by in sb.items():
        t_ms = stages.get(st, 0.0)
        if t_ms > 0:
            stage_roofline[st] = {
                "ms": t_ms,
                "bytes_definition": by["definition"], "bytes_design": by["design"],
                "frac_definition": round(by["definition"] / (t_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                "frac_design": round(by["design"] / (t_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    # the kernel with the largest total time among those with an algorithmic
    # byte count (under a counter-collecting profiler the order can change)
    dom = max((k for k in kt_all if kernel_bytes(k, n, C, V, A, P, M, B)),
              key=lambda k: kt_all[k][1])

    # timed region: K steps, events around the dominant kernel's launches only
    ctx.timing_reset()
    ctx.set_timing(True, only=dom)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reconstruct_plain(g, holos, mask, re, im, pos)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    kt = ctx.timing()
    ms = (t1 - t0) * 1e3 / args.steps
    print(f"[bench] {args.steps} steps, {ms:.3f} ms/volume", file=sys.stderr, flush=True)

    pcie = None
    if not args.no_pcie:
        # PCIe-inclusive rate: the holograms start in pinned host memory and the
        # library streams them to HBM inside each timed step, each chunk's copy
        # overlapped with the previous chunk's transform (the executables' case)
        host = torch.empty(holos.shape, dtype=holos.dtype, pin_memory=True)
        host.copy_(holos)
        ctx.reconstruct_plain(g, host, mask, re, im, pos)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.reconstruct_plain(g, host, mask, re, im, pos)
        torch.cuda.synchronize()
        ms_pcie = (time.perf_counter() - t0) * 1e3 / args.steps
        del host
        pcie = {"value": round(A / (ms_pcie / 1e3), 2), "unit": "holograms/s",
                "ms_per_step": round(ms_pcie, 4),
                "note": "pinned host holograms streamed to HBM by the library inside the step "
                        "(copies overlapped with the hologram stage)"}

    # roofline of the dominant kernel, from the timed region's events
    launches, tot = kt[dom]
    lps = launches / args.steps
    bpl = kernel_bytes(dom, n, C, V, A, P, M, B) / lps
    avg_s = tot / launches * 1e-3
    achieved = bpl / avg_s / 1e9
    traffic = counters.get(dom) if isinstance(counters.get(dom), (int, float)) else None
    step_def = sum(v["definition"] for v in sb.values())
    step_des = sum(v["design"] for v in sb.values())
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic, "algorithmic_bytes_per_launch": int(bpl),
                "ms_per_launch": round(avg_s * 1e3, 5),
                "stages": stage_roofline,
                "step": {"bytes_definition": step_def, "bytes_design": step_des,
                         "frac_definition": round(step_def / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                         "frac_design": round(step_des / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
                "note": ("achieved = the kernel's share of the per-hologram stage bytes "
                         "(hologram rows: the b_in N^2 hologram read) per launch over its "
                         "event-timed mean duration; stages and step: their bytes over their "
                         "time")}

    cpu = None
    if not args.no_cpu and args.workload == "cfg2":
        runs, nproc, model = cpu_baseline(holos, mask, A)
        ref4 = runs["NB_THREADS_4"]
        cpu = {"value": ref4["holograms_per_s"], "unit": "holograms/s", "cores": 4,
               "kind": "port",
               "sample": (f"all {A} holograms of the workload, the reference's FP64 "
                          f"arithmetic at its stage boundaries (oracle/cpu_pipeline.py), "
                          f"FFTs on NB_THREADS=4 threads as the shipped recon.txt, host "
                          f"loops serial; {ref4['fft']}"),
               "ms_per_volume": ref4["ms_per_volume"],
               "nproc": nproc, "cpu_model": model, "settings": runs,
               "gpu_speedup": {k: round(r["ms_per_volume"] / ms, 1) for k, r in runs.items()}}

    hbm = copy_bandwidth(torch)
    if args.no_quality:
        quality = None
    elif args.workload == "large":  # 50 iterations at 1024^3 on the large acquisition
        quality = adjoint_quality(ctx, torch, np, m, synth, "large", 50, 50, graph=True)
    else:
        quality = adjoint_quality(ctx, torch, np, m, synth)

    out = {
        "metric": METRIC, "value": round(A / (ms / 1e3), 2), "unit": "holograms/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain method: {A} x {n}^2 uint16 holograms -> crop {C}^2 -> "
                               f"count-normalised {V}^3 complex64 spectrum (only its hit band "
                               f"materialised; counts not stored) -> {V}^3 float32 index and "
                               f"absorption",
                   "holograms": A, "hologram_side": n, "crop": C, "volume": V,
                   "hit_band": [lo, hi]},
        "ms_per_volume": round(ms, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "hbm_copy": hbm,
        "adjoint_tv": quality,
        "kernels": per_kernel,
        "stages_ms": stages,
        "holograms_per_s_s1_to_s3": round(A / ((ms - stages.get("S4", 0.0)) / 1e3), 2),
        "kernels_note": "HIP-event durations per kernel from K untimed steps with every "
                        "kernel timed; the timed region times only the roofline kernel",
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
