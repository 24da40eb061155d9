"""Benchmark: batched forward + adjoint multipoles on one MI355X.

Workload (the batched configuration): 256 seeded P(k) variants with
A ~ U[1.5e4, 2.5e4], tilt ~ U[0.92, 1.0], BAO scale ~ U[95, 115] and ragged k-grid
lengths from {1024, ..., 16384} (FFT length = grid length), flat-sky density+RSD
multipoles l = 0, 2, 4 at z_mean = 1 (b = 1.5, f and D1 from LCDM with
Omega_m = 0.31), 512 separations linearly spaced in [5, 300] Mpc/h.  One step =
forward multipoles of the whole batch plus the vector-Jacobian product with respect
to every P(k) sample (prepare -> tables(adjoint) -> project -> prepare_adj).

Metric: (l, r) multipole points per second, counting only points within 1e-8
relative of the oracle (CPU restatement of the reference algorithm).  The plain
per-call GPU path (one spectrum and one table per twofast_1bessel call, host
buffers, host Akima and projection) and the reference-algorithm CPU path (oracle,
OpenMP) are timed on bounded samples beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4)]
LS = [0, 2, 4]
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def build_inputs(nspec=256, nsep=512, seed=0):
    from tests.synth import batch_spectra, flat_weights
    from coffe_amd import background as B
    specs = batch_spectra(nspec, seed=seed)
    sep = np.linspace(5.0, 300.0, nsep) / H
    bg = B.compute(omega_cdm=0.26, omega_baryon=0.05, h=H)
    zbar = 1.0
    from coffe_amd.engine import akima
    f = float(akima(bg["z"], bg["f"], [zbar])[0])
    D1 = float(akima(bg["z"], bg["D1"], [zbar])[0])
    A = flat_weights(LS, 1.5, f, D1, nsep)
    return specs, sep, A, bg, dict(f=f, D1=D1, b=1.5, zbar=zbar)


def algorithmic_flops(n_fft, n_in, nsep, ntab, nq=1):
    """algorithmic flops per kernel launch (real FFT of length N = 2.5 N log2 N; complex
    multiply 6; natural-spline system build + solve 12 per node; spline evaluation 12 per
    resample point; Akima 40 per value and 80 more for its VJP)."""
    n_fft = np.asarray(n_fft, dtype=np.float64)
    n_in = np.asarray(n_in, dtype=np.float64)
    fft = 2.5 * n_fft * np.log2(n_fft)
    M1 = n_fft / 2 + 1
    return dict(
        k_resample=float(np.sum(12 * n_in + 12 * n_fft)),
        k_prepare=float(nq * np.sum(n_fft + fft + 8 * M1)),
        k_tables=float(ntab * np.sum(2 * 6 * M1 + 2 * fft + 2 * 8 * M1 + 120.0 * nsep)),
        k_project=float(2.0 * ntab * 3 * nsep * len(n_fft)),
        k_adj_c2r=float(nq * np.sum(fft + 8 * M1 + 2 * ntab * M1 + n_fft)),
        k_spline_adj=float(np.sum(16 * n_fft + 12 * n_in + 10 * n_in)),
    )


def hbm_bytes(n_fft, n_in, nsep, ntab, nspec, nz, nl, nq=1):
    """compulsory HBM bytes per kernel launch: inputs read once, grid tables read once
    per work unit, workspaces written once and read once, outputs written once."""
    n_fft = np.asarray(n_fft, dtype=np.float64)
    n_in = np.asarray(n_in, dtype=np.float64)
    M1 = n_fft / 2 + 1
    return dict(
        k_resample=float(np.sum(16 * n_in + 8 * n_fft + 8 * n_fft)),
        k_prepare=float(nq * np.sum(8 * n_fft + 8 * n_fft + 16 * M1)),
        k_tables=float(ntab * np.sum(16 * M1 * 3 + 8 * 6 * 2 * nsep + 16 * M1) + 8 * nspec * ntab * nsep +
                       8 * nspec * nz * nl * nsep),
        k_project=float(8 * nspec * ntab * nsep + 8 * nz * nl * ntab * nsep + 8 * nspec * nz * nl * nsep),
        k_adj_c2r=float(np.sum(ntab * 16 * M1 + 8 * n_fft * 2)),
        k_spline_adj=float(np.sum(8 * n_in + 8 * n_fft * 2 + 8 * n_in)),
    )


def cpu_baseline(specs, sep, bg, phys, sample):
    import oracle as O
    sub = specs[:sample]
    bgo = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"])
    zb = np.arange(11.0)
    b = np.full(11, phys["b"])
    t0 = time.perf_counter()
    O.reference_batch([(s[0], s[1]) for s in sub], [len(s[0]) for s in sub], bgo, zb, b, b, [phys["zbar"]], LS, sep,
                      den=True, rsd=True, d1=False, flatsky=True, small_r=True)
    dt = time.perf_counter() - t0
    points = len(sub) * len(LS) * len(sep)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=points / dt, unit="points/s", cores=cores, kind="port",
                sample=f"{len(sub)} spectra x {len(LS)} l x {len(sep)} seps, reference sequence (8 FFTLog tables with "
                       f"small-r QAG nodes, QAG-GK61 mu-integral per point), forward only, {dt:.2f} s wall")


def plain_path(specs, sep, A, sample):
    """per-call GPU FFTLog: one spectrum and one table per twofast_1bessel call (host buffers,
    synchronous), then host Akima at the separations and host projection."""
    from coffe_amd.engine import akima, twofast_1bessel
    sub = specs[:sample]
    t0 = time.perf_counter()
    for k, p, _ in sub:
        N = len(k)
        itab = np.empty((len(TABLES), len(sep)))
        for t, (n, l) in enumerate(TABLES):
            rx, ry = twofast_1bessel(k, p, N, l, n)
            itab[t] = akima(rx, ry, sep)
        np.einsum("zltr,tr->zlr", A, itab)
    dt = time.perf_counter() - t0
    points = len(sub) * len(LS) * len(sep)
    return dict(value=points / dt, unit="points/s", sample=f"{len(sub)} spectra, forward only, {dt:.3f} s wall")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nspec", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (5 times), to scale HBM counters")
    args = ap.parse_args()
    if args.copy_calibration_mb:
        import torch
        n = args.copy_calibration_mb * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the engine is single-device (spectra shard across GPUs with no exchange; "
              "see DESIGN.md). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from coffe_amd.engine import Batch, Engine
    import oracle as O

    specs, sep, A_np, bg, phys = build_inputs(args.nspec)
    nspec, nsep = len(specs), len(sep)
    n_in = [len(s[0]) for s in specs]
    n_fft = n_in
    eng = Engine(0)
    batch = Batch(eng, n_in, n_fft, [s[0][0] for s in specs], [s[0][-1] for s in specs], TABLES, sep, 1, len(LS))
    dev = "cuda"
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device=dev)
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device=dev)
    A = torch.tensor(A_np, device=dev)
    rng = np.random.default_rng(1)
    gout = torch.tensor(rng.standard_normal((nspec, 1, len(LS), nsep)), device=dev)
    itab = torch.empty((nspec, len(TABLES), nsep), dtype=torch.float64, device=dev)
    out = torch.empty((nspec, 1, len(LS), nsep), dtype=torch.float64, device=dev)
    gpk = torch.empty_like(p)
    stream = torch.cuda.current_stream()
    sp = __import__("ctypes").c_void_p(stream.cuda_stream)
    kernels = ["k_resample", "k_prepare", "k_tables", "k_project", "k_adj_c2r", "k_spline_adj"]
    launches = [lambda: batch.kernel_resample(k, p, sp), lambda: batch.kernel_fft_prepare(sp),
                lambda: batch.stage_tables(A, gout, itab, sp), lambda: batch.stage_project(A, itab, out, sp),
                lambda: batch.kernel_adj_c2r(sp), lambda: batch.kernel_spline_adj(k, gpk, sp)]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    kernel_ms = {kk: 0.0 for kk in kernels}

    def step():
        # one forward + adjoint pass over the batch: a single C call launches the six
        # kernels back to back on the stream (resample, prepare, tables, project,
        # adj_c2r, spline_adj)
        batch.vjp(k, p, A, gout, itab=itab, out=out, gpk=gpk, stream=sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps

    # per-kernel durations: each kernel launched `reps` times back to back between two
    # events on the launch stream (the kernels are idempotent given their inputs, and
    # the workspaces hold the state of a full step), right after the timed region
    reps = 10
    for i, launch in enumerate(launches):
        launch()
        torch.cuda.synchronize()
        ev[0].record(stream)
        for _ in range(reps):
            launch()
        ev[1].record(stream)
        torch.cuda.synchronize()
        kernel_ms[kernels[i]] = ev[0].elapsed_time(ev[1]) / reps

    # parity: every point against the oracle (tables at the separations, flat-sky weights)
    ref_tab = O.tables_at_sep_batch([(s[0], s[1]) for s in specs], n_fft, TABLES, sep)
    ref = np.einsum("zltr,str->szlr", A_np, ref_tab)
    got = out.cpu().numpy()
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    ok = rel <= 1e-8
    points = nspec * len(LS) * nsep
    good_points = int(ok.sum())
    value = good_points / (ms_per_step / 1e3)

    flops = algorithmic_flops(n_fft, n_in, nsep, len(TABLES))
    bytes_ = hbm_bytes(n_fft, n_in, nsep, len(TABLES), nspec, 1, len(LS))
    dom = max(kernels, key=lambda kk: kernel_ms[kk])
    achieved = flops[dom] / (kernel_ms[dom] / 1e3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_c3.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom)
    res = {
        "metric": "(l,r) multipole points/sec, batched fwd+adjoint vs plain FFTLog, 1 MI355X @1e-8 rel",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched: 256 seeded P(k), ragged N in {1024..16384}, 512 seps 5-300 Mpc/h, "
                               "l=0,2,4 flat-sky den+RSD z=1, forward+adjoint",
                   "spectra": nspec, "separations": nsep, "multipoles": LS, "tables": TABLES,
                   "points_per_step": points, "fft_lengths": sorted(set(int(n) for n in n_fft)),
                   "threads_per_wg": int(os.environ.get("COFFE_ENGINE_THREADS", 512))},
        "roofline": {"bound": "mfma", "kernel": dom,
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "note": "FP64 VALU bound (no MFMA on this path); peak = MI355X FP64 78.6 TFLOP/s; "
                             "traffic = HBM bytes/launch from rocprofv3 FETCH_SIZE+WRITE_SIZE scaled by a copy kernel",
                     "kernel_us": {kk: kernel_ms[kk] * 1e3 for kk in kernels},
                     "kernel_flops": flops, "kernel_algorithmic_bytes": bytes_,
                     "kernel_fp64_frac": {kk: flops[kk] / (kernel_ms[kk] / 1e3) / PEAK_FP64 for kk in kernels},
                     "kernel_hbm_frac": {kk: bytes_[kk] / (kernel_ms[kk] / 1e3) / PEAK_HBM for kk in kernels}},
        "parity": {"points": points, "within_1e-8": good_points, "max_rel": float(rel.max()),
                   "median_rel": float(np.median(rel))},
    }
    if args.plain_sample > 0:
        pp = plain_path(specs, sep, A_np, args.plain_sample)
        res["plain_path"] = pp
        res["vs_plain"] = value / pp["value"]
    if not args.no_cpu:
        cores = os.cpu_count() or 1
        sample = args.cpu_sample or nspec
        res["cpu_baseline"] = cpu_baseline(specs, sep, bg, phys, sample)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
