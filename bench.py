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


def traffic_of(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of `workload` (None when absent)"""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(tfile):
        return None
    return json.load(open(tfile)).get(kernel)


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


def cpu_baseline(specs, sep, bg, phys, sample, runs=5, sample_1t=None):
    """the oracle's reference sequence on the host: the median of `runs` timings at the
    OpenMP thread count of the box (OMP_NUM_THREADS) on `sample` spectra, and the median
    of `runs` one-thread timings on `sample_1t` spectra (default sample / 16)"""
    import oracle as O
    bgo = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"])
    zb = np.arange(11.0)
    b = np.full(11, phys["b"])

    def rate(sub):
        t0 = time.perf_counter()
        O.reference_batch([(s[0], s[1]) for s in sub], [len(s[0]) for s in sub], bgo, zb, b, b, [phys["zbar"]], LS,
                          sep, den=True, rsd=True, d1=False, flatsky=True, small_r=True)
        dt = time.perf_counter() - t0
        return len(sub) * len(LS) * len(sep) / dt, dt

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    prev = O.set_threads(cores)
    try:
        sub = specs[:sample]
        multi = sorted(rate(sub) for _ in range(runs))
        sub1 = specs[:sample_1t or max(1, sample // 16)]
        O.set_threads(1)
        single = sorted(rate(sub1) for _ in range(runs))
    finally:
        O.set_threads(prev)
    med, med1 = multi[runs // 2], single[runs // 2]
    return dict(value=med[0], unit="points/s", cores=cores, kind="port",
                sample=f"{len(sub)} spectra x {len(LS)} l x {len(sep)} seps, reference sequence (8 FFTLog tables with "
                       f"small-r QAG nodes, QAG-GK61 mu-integral per point), forward only; median of {runs} runs "
                       f"({med[1]:.2f} s wall each) at {cores} threads",
                runs=[r[0] for r in multi],
                one_thread={"value": med1[0], "unit": "points/s", "cores": 1,
                            "sample": f"{len(sub1)} spectra, median of {runs} runs ({med1[1]:.2f} s wall each)",
                            "runs": [r[0] for r in single]})


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


def run_fullsky(args):
    """Full-sky density + RSD + Doppler-1 multipoles of one P(k) (N = 16384 input and
    FFT samples, the 8 default tables) at 8 mean redshifts x l = 0, 2, 4 x 512
    separations, mu integral by 256-node Gauss-Legendre on the GPU (k_mu_weights)."""
    import torch
    import oracle as O
    from coffe_amd import background as B
    from coffe_amd.engine import Batch, Engine, mu_weights
    from tests.synth import spectrum
    tabs = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]
    k, pk = spectrum(16384)
    sep = np.linspace(5.0, 300.0, 512) / H
    zs = np.linspace(0.5, 2.0, 8)
    bgd = B.compute(omega_cdm=0.26, omega_baryon=0.05, h=H)
    bg = dict(z=bgd["z"], chi=bgd["chi"], D1=bgd["D1"], f=bgd["f"], cH=bgd["conformal_Hz"], G1=bgd["G1"],
              G2=bgd["G2"])
    zb, b = np.linspace(0, 10, 11), np.full(11, 1.5)
    eng = Engine(0)
    batch = Batch(eng, [len(k)], [len(k)], [k[0]], [k[-1]], tabs, sep, len(zs), len(LS))
    kt, pt = torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda")
    A = torch.empty((len(zs), len(LS), len(tabs), len(sep)), dtype=torch.float64, device="cuda")

    def step():
        mu_weights(eng, bg, zb, b, b, zs, LS, tabs, sep, den=1, rsd=1, d1=1, flatsky=False, nmu=256, out=A)
        return batch.forward(kt, pt, A)[1]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the mu-weight launch timed on its stream: 10 calls back to back between two HIP
    # events (each call = one upload of the inputs, k_akima_coef, k_mu_weights)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = 10
    ev[0].record(stream)
    for _ in range(reps):
        mu_weights(eng, bg, zb, b, b, zs, LS, tabs, sep, den=1, rsd=1, d1=1, flatsky=False, nmu=256, out=A)
    ev[1].record(stream)
    torch.cuda.synchronize()
    mu_us = ev[0].elapsed_time(ev[1]) * 1e3 / reps
    mu_flops = 600.0 * len(zs) * len(LS) * len(sep) * 256
    got = out.cpu().numpy()[0]
    otabs = {t: O.integral_table(t[0], t[1], k, pk, len(k), small_r=False) for t in tabs}
    # the truth: the reference's formulas in binary128 (the oracle's high-precision
    # weights) times the oracle's tables; every point counts, zero crossings included
    hp = O.mu_weights_hp(bg, zb, b, b, zs, LS, sep, den=1, rsd=1, d1=1, flatsky=False, nmu=256)
    It = np.array([O.akima(otabs[t][0], otabs[t][1], sep) for t in tabs])
    ref = np.einsum("zltr,tr->zlr", hp, It)
    rel = np.abs(got - ref) / np.abs(ref)
    points = got.size
    good = int((rel <= 1e-8).sum())
    # the FP64 evaluation of the reference's own form (GL-256) for comparison
    ref64 = O.multipoles(otabs, bg, zb, b, b, zs, LS, sep, den=1, rsd=1, d1=1, flatsky=False, quad=256)
    good64 = int((np.abs(ref64 - ref) / np.abs(ref) <= 1e-8).sum())
    t1 = time.perf_counter()
    sub = sep[::32]
    O.multipoles(otabs, bg, zb, b, b, zs[:2], LS, sub, den=1, rsd=1, d1=1, flatsky=False, quad=0, epsrel=1e-5)
    cpu_pts = 2 * len(LS) * len(sub) / (time.perf_counter() - t1)
    print(json.dumps({
        "metric": "full-sky multipole points/sec (8 z x l=0,2,4 x 512 r, 8 tables, GL-256 mu), 1 MI355X",
        "value": points / (ms / 1e3), "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "full-sky den+RSD+d1, one P(k) with N=16384, 8 tables, z=linspace(0.5,2,8), "
                               "512 seps 5-300 Mpc/h, GL-256; step = mu weights + tables + projection"},
        "roofline": {"bound": "mfma", "kernel": "k_mu_weights", "achieved": mu_flops / (mu_us * 1e-6) / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": mu_flops / (mu_us * 1e-6) / PEAK_FP64,
                     "traffic": traffic_of("fullsky_8z_512sep", "k_mu_weights"), "launch_us": mu_us,
                     "note": "600 flop per (z, l, r, mu node) x 8 z x 3 l x 512 r x 256 nodes per launch; FP64 VALU "
                             "bound (no matrix work), peak = MI355X FP64 78.6 TFLOP/s; launch time from HIP events "
                             "over 10 back-to-back calls, each including the input upload and k_akima_coef"},
        "parity": {"points": points, "within_1e-8": good, "max_rel": float(rel.max()),
                   "fp64_reference_form_within_1e-8": good64,
                   "note": "against the reference's formulas in binary128 under the same GL-256 rule (oracle "
                           "mu_weights_hp) times the oracle's tables; fp64_reference_form_within_1e-8 counts the "
                           "points where an FP64 evaluation of the reference's own form meets the same bar"},
        "cpu_baseline": {"value": cpu_pts, "unit": "points/s", "cores": 1, "kind": "port",
                         "sample": f"2 z x 3 l x {len(sub)} seps, QAG-GK61 mu integral per point on the oracle tables "
                                   "(tables not counted)"}}))


def run_covariance(args):
    """2D FFTLog Gaussian covariance: 64 P(k), N = 8000, l = 0, 2, 4, 256 separations
    10-200 Mpc/h, 4 redshift bins (integrals of P and P^2 for the 6 pairs l1 <= l2, then
    the assembly)."""
    import torch
    import oracle as O
    from coffe_amd.covariance import CovPlan
    from coffe_amd.engine import Engine
    from tests.synth import batch_spectra
    specs = batch_spectra(64, seed=5, lengths=(4096,))
    sep = np.linspace(10.0, 200.0, 256) / H
    kmin = max(s[0][0] for s in specs)
    kmax = min(s[0][-1] for s in specs)
    eng = Engine(0)
    plan = CovPlan(eng, 8000, kmin, kmax, LS, sep)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    n_in = [len(s[0]) for s in specs]
    nz = 4
    zargs = dict(f=[0.7, 0.8, 0.85, 0.9], D1=[0.8, 0.7, 0.6, 0.5], nbar=[1e-3 * H ** 3] * nz,
                 vol=[1e9 / H ** 3] * nz, pix=[5.0 / H] * nz)

    def step():
        integ = plan.integrals(k, p, n_in)
        return plan.assemble(integ, [1.5] * nz, **zargs), integ

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cov, integ = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    entries = cov.numel()
    # the integrals launch (k_cov_diag, k_cov_rows, k_cov_cols) timed on its stream with
    # HIP events over 3 calls back to back; k_cov_rows is the dominant kernel
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = 3
    ev[0].record(stream)
    for _ in range(reps):
        plan.integrals(k, p, n_in, out=integ)
    ev[1].record(stream)
    torch.cuda.synchronize()
    rows_us = ev[0].elapsed_time(ev[1]) * 1e3 / reps
    # the adjoint of the integrals (d<G, integrals>/dP, coffe_cov_integrals_vjp) on the
    # same batch, a random cotangent; the call synchronises its stream
    gint = torch.randn(integ.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    gpk = plan.integrals_vjp(k, p, n_in, gint)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(reps):
        plan.integrals_vjp(k, p, n_in, gint, gpk=gpk)
    ev[1].record(stream)
    torch.cuda.synchronize()
    vjp_us = ev[0].elapsed_time(ev[1]) * 1e3 / reps
    N = 8000
    row_flops = 7.5 * N * np.log2(N) + 8.0 * N + 2.0 * 72 * len(sep)
    nrows_launch = 64 * 2 * plan.npair * plan.nrows
    cov_flops = row_flops * nrows_launch
    res = {
        "metric": "Gaussian covariance entries/sec (64 P(k), 2D FFTLog N=8000, 4 z x (3 l x 256 r)^2), 1 MI355X",
        "value": entries / (ms / 1e3), "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 seeded P(k) (4096 samples), 2D FFTLog N=8000, l=0,2,4, 256 seps 10-200 Mpc/h, "
                               "4 z bins; step = P and P^2 integrals of 6 pairs + assembly",
                   "transform_rows_per_pair": plan.nrows, "transforms_per_step": 64 * 2 * plan.npair},
        "roofline": {"bound": "mfma", "kernel": "k_cov_rows", "achieved": cov_flops / (rows_us * 1e-6) / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": cov_flops / (rows_us * 1e-6) / PEAK_FP64,
                     "traffic": traffic_of("covariance_64spec_n8000", "k_cov_rows"), "launch_us": rows_us,
                     "note": "per transform row 7.5 N log2 N (complex FFT_N + c2r) + 8 N (kernel products) + "
                             "2 x 72 x R (column functionals), N = 8000, R = 256, x 64 spectra x 2 powers x 6 pairs x "
                             "rows per pair; time = HIP events over 3 integrals calls (k_cov_diag + k_cov_rows + "
                             "k_cov_cols, rows > 95 % in the kernel trace); FP64 peak 78.6 TFLOP/s"},
        "adjoint": {"integrals_us": rows_us, "integrals_vjp_us": vjp_us,
                    "note": "coffe_cov_integrals_vjp on the same 64 spectra (d<G, integrals>/dP for every input "
                            "sample), HIP events over 3 calls"},
    }
    if args.cpu_sample:
        ks, ps = specs[0][0], specs[0][1]
        t1 = time.perf_counter()
        ref = O.covariance_fftlog_integral(ks, ps, 1.0, 0, 2, 8000, sep, kmin, kmax)
        dt = time.perf_counter() - t1
        g = integ.cpu().numpy()[0, 0, 1] / (5 * 2.0 / np.pi ** 2)
        res["parity"] = {"checked": "spectrum 0, P, (l1, l2) = (0, 2), 256 x 256 against the oracle",
                         "max_abs_over_max": float(np.max(np.abs(g - ref)) / np.max(np.abs(ref)))}
        res["cpu_baseline"] = {"value": entries / (64 * 2 * plan.npair) / dt, "unit": "entries/s", "cores": 1,
                               "kind": "port", "sample": f"one N=8000 2D FFTLog + bicubic of the oracle, {dt:.1f} s, "
                                                         "scaled to the entries of one transform"}
    print(json.dumps(res))
    plan.close()


DIV_FLOP_PER_EVAL = 145.0


def run_divergent(args):
    """The divergent table I_0^4 of the reference's configuration (integrals.c:772-909):
    the CLASS spectrum of the test suite, k in [1e-5, 300], the 297 separations of
    integrals.c:76-99 (I_0^4 and the r = 0 renormalisation) and the 201 x 201
    renormalisation grid on [0, chi_max], chi_max = 2477.9 Mpc (as in
    benchmark_integral8_renormalization.dat): 41,005 QAG-GK61 integrals per step, one
    wavefront each (k_divergent)."""
    import ctypes
    import torch
    import oracle as O
    from coffe_amd._lib import check, load_library
    from coffe_amd.engine import COFFE_SEP, Engine
    k, pk = np.loadtxt(os.path.join(ROOT, "tests", "golden", "PkL_CLASS.dat")).T
    kmin, kmax = 1e-5, 300.0
    chi = np.linspace(0.0, 0.8265275 * 2997.92458, 201)
    sep = COFFE_SEP
    units = 2 * len(sep) + len(chi) ** 2
    eng = Engine(0)
    L = load_library()
    dev = dict(device="cuda", dtype=torch.float64)
    kt, pt = torch.tensor(k, **dev), torch.tensor(pk, **dev)
    st, ct = torch.tensor(sep, **dev), torch.tensor(chi, **dev)
    i40, r0 = torch.empty(len(sep), **dev), torch.empty(len(sep), **dev)
    ren = torch.empty((len(chi), len(chi)), **dev)
    status = torch.empty(units, dtype=torch.int32, device="cuda")
    nint = torch.empty(units, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        check(L.coffe_divergent_tables_device(eng.h, kt.data_ptr(), pt.data_ptr(), len(k), kmin, kmax,
                                              st.data_ptr(), len(sep), ct.data_ptr(), len(chi), i40.data_ptr(),
                                              r0.data_ptr(), ren.data_ptr(), status.data_ptr(), nint.data_ptr(),
                                              ctypes.c_void_p(stream.cuda_stream)), "coffe_divergent_tables_device")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the kernel alone: HIP events around back-to-back launches on the launch stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = 10
    ev[0].record(stream)
    for _ in range(reps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    us = ev[0].elapsed_time(ev[1]) * 1e3 / reps
    n = nint.cpu().numpy().astype(np.float64)
    evals = float(np.sum(61.0 * (2.0 * n - 1.0)))
    flops = DIV_FLOP_PER_EVAL * evals
    stat = status.cpu().numpy()
    x, y = np.loadtxt(os.path.join(ROOT, "tests", "golden", "benchmark_integral8.dat")).T
    rel = np.abs(i40.cpu().numpy() * (1 / 2997.92458) ** 4 / y - 1)
    x1, x2, v = np.loadtxt(os.path.join(ROOT, "tests", "golden", "benchmark_integral8_renormalization_every10.dat")).T
    g = ren.cpu().numpy()[::10, ::10].reshape(-1) * (1 / 2997.92458) ** 4
    rel2 = np.abs(g / v - 1)
    res = {
        "metric": "divergent-table integrals/sec (I_0^4 at 297 seps + r=0 renormalisation + 201x201 grid), 1 MI355X",
        "value": units / (ms / 1e3), "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "reference test spectrum (PkL_CLASS.dat)",
        "config": {"workload": "I_0^4 and its renormalisations of the reference's test configuration: k in [1e-5, 300], "
                               "297 separations, 201 x 201 chi grid to 2477.9 Mpc; QAG-GK61 epsrel 1e-5 per integral",
                   "integrals_per_step": units},
        "roofline": {"bound": "mfma", "kernel": "k_divergent", "achieved": flops / (us * 1e-6) / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": flops / (us * 1e-6) / PEAK_FP64,
                     "traffic": traffic_of("divergent_ref_config", "k_divergent"), "launch_us": us,
                     "note": f"{DIV_FLOP_PER_EVAL:.0f} flop per integrand evaluation (Akima 50, two j_0 with sincos "
                             "80, rule sums 15), 61 (2 n - 1) evaluations per integral of n final intervals (counted "
                             f"on the device: {evals:.3e} per launch); one wavefront per integral, its bisections "
                             "serial: latency-bound; FP64 peak 78.6 TFLOP/s"},
        "parity": {"i40_vs_fixture_max_rel": float(rel.max()), "grid_vs_fixture_max_rel": float(rel2.max()),
                   "interval_cap_hits": int(np.sum(stat == 1)),
                   "note": "benchmark_integral8.dat (11 digits) and every 10th node of "
                           "benchmark_integral8_renormalization.dat (7 digits); the one cap hit is the r = 0 "
                           "renormalisation at sep = 1e-6 Mpc, which the reference discards (integrals.c:838-839)"},
    }
    if not args.no_cpu:
        t1 = time.perf_counter()
        O.divergent_tables(k, pk, sep, chi[::2], kmin=kmin, kmax=kmax)
        dt = time.perf_counter() - t1
        n_cpu = 2 * len(sep) + len(chi[::2]) ** 2
        res["cpu_baseline"] = {"value": n_cpu / dt, "unit": "integrals/s", "cores": os.cpu_count(), "kind": "port",
                               "sample": f"the oracle (C, OpenMP over the integrals) on the 297 separations and a "
                                         f"101 x 101 grid ({n_cpu} integrals, {dt:.2f} s)"}
    print(json.dumps(res))


# FP64 operations per integrand evaluation of the integrated kernels with every contribution on,
# counted from the source: ~19 Akima lookups of ~40 flop (five slopes and the cubic; the
# bisections are comparisons), the bicubic renormalisation ~60, the terms ~200
INTG_FLOP_PER_NODE = 1000.0


def run_integrated(args):
    """The integrated terms of the reference's all-contributions configuration (density,
    RSD, Doppler, d2, g1-g3, lensing, g4, g5, full sky; functions.c:1486-4017): the
    single- and double-integrated correlation function at z = 1 on the 64 theta nodes of
    the multipoles and 64 separations 5-300 Mpc/h, projected on l = 0, 2, 4
    (coffe_integrated_corrfunc, one workgroup of k_integrated_single and one of
    k_integrated_double per (mu, separation))."""
    import torch
    import oracle as O
    from oracle import integrated as OI
    from coffe_amd import background as B
    from coffe_amd.engine import COFFE_SEP, Engine, integrated_corrfunc
    H = 0.67
    k, pk = np.loadtxt(os.path.join(ROOT, "tests", "golden", "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    Engine(0)
    tabs = {key: O.integral_table(key[0], key[1], k, pk, 10000, small_r=True) for key in O.TABLE_ORDER}
    chi = np.arange(201) / 200 * O.akima(bg["z"], bg["chi"], [1.2])[0]
    i40, r0, ren = O.divergent_tables(k, pk, COFFE_SEP, chi)
    tabs[(4, 0)] = (COFFE_SEP, i40)
    x0, y0 = COFFE_SEP.copy(), r0.copy()
    x0[0], y0[0] = 0.0, 0.0
    zb, one, zero = np.arange(11.0), np.ones(11), np.zeros(11)
    bgd = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], a=bg["a"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"],
               G2=bg["G2"])
    flags = ["den", "rsd", "d1", "d2", "g1", "g2", "g3", "len", "g4", "g5"]
    t_, w_ = np.polynomial.legendre.leggauss(64)
    theta = np.pi / 2 * (t_ + 1)
    mus, wm = np.cos(theta), np.pi / 2 * w_ * np.sin(theta)
    seps = np.linspace(5, 300, 64) / H
    ls = [0, 2, 4]
    pls = np.array([np.polynomial.legendre.legval(mus, [0] * l + [1]) * (2 * l + 1) / 2 for l in ls])
    args_c = (bgd, zb, one, one, zb, zero, zero, zero, zero, 0.30, H / 2997.92458, tabs, flags)

    def step():
        s_, d_ = integrated_corrfunc(*args_c, [1.0], mus, seps, ren=(chi, ren), ren0=(x0, y0))
        return np.einsum("lm,m,mr->lr", pls, wm, (s_ + d_)[0])

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    points = len(ls) * len(seps)
    # nodes per (mu, r): the double rule (8 or 12 x 16 x1 nodes x 192 u nodes) and the single
    # rule (up to 4 pieces x 512): counted from the cut positions of each point
    nodes = 0
    for m in mus:
        for r in seps:
            cm = float(O.akima(bg["z"], bg["chi"], [1.0])[0])
            chi1, chi2 = cm - r * m / 2, cm + r * m / 2
            c = (2 * cm * cm - r * r + m * m * r * r / 2) / (2 * cm * cm - m * m * r * r / 2)
            nodes += (12 if chi1 * c > chi2 else 8) * 16 * 192
            inside = {round(v, 15) for v in (chi1 * c / chi2, chi2 * c / chi1) if 0 < v < 1}
            nodes += (1 + (2 * len(inside) if len(inside) == 2 else len(inside))) * 512
    flops = INTG_FLOP_PER_NODE * nodes
    # parity: three points against the oracle on the same rules
    chk = []
    for im, ir in ((5, 3), (32, 20), (60, 50)):
        fl = {f_: 1 for f_ in flags}
        inp = OI.Inputs(tabs, bg, 0.30, H / 2997.92458, zb, one, one, zb, zero, zero, zero, zero,
                        div=(COFFE_SEP, i40, r0), ren=(chi, ren))
        s_, d_ = integrated_corrfunc(*args_c, [1.0], [mus[im]], [seps[ir]], ren=(chi, ren), ren0=(x0, y0))
        want = OI.corrfunc_single(inp, fl, 1.0, mus[im], seps[ir]) + OI.corrfunc_double(inp, fl, 1.0, mus[im], seps[ir])
        chk.append(abs(float(s_[0, 0, 0] + d_[0, 0, 0]) / want - 1))
    res = {
        "metric": "integrated-term multipole points/sec (all contributions, full sky), 1 MI355X",
        "value": points / (ms / 1e3), "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "reference test spectrum (PkL_CLASS.dat)",
        "config": {"workload": "single- and double-integrated correlation function, every contribution, z = 1, "
                               "64 theta nodes x 64 separations 5-300 Mpc/h, projected on l = 0, 2, 4",
                   "points_per_step": points, "integrand_nodes_per_step": nodes},
        "roofline": {"bound": "mfma", "kernel": "k_integrated_single + k_integrated_double", "unit": "TFLOP/s",
                     "achieved": flops / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "frac": flops / (ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "note": f"{INTG_FLOP_PER_NODE:.0f} flop per integrand node (estimated from the source), "
                             "divided by the whole synchronous call (uploads, allocation, the kernel, download): "
                             "a lower bound of the kernels' own rate, whose durations the rocprof summary gives; "
                             "FP64 VALU, with table lookups bound by the L1's gather throughput (random lines per lane)"},
        "parity": {"vs_oracle_max_rel": max(chk), "points": len(chk)},
    }
    if not args.no_cpu:
        inp = OI.Inputs(tabs, bg, 0.30, H / 2997.92458, zb, one, one, zb, zero, zero, zero, zero,
                        div=(COFFE_SEP, i40, r0), ren=(chi, ren))
        fl = {f_: 1 for f_ in flags}
        t1 = time.perf_counter()
        n = 0
        for m in mus[::8]:
            OI.corrfunc_single(inp, fl, 1.0, m, seps[10])
            OI.corrfunc_double(inp, fl, 1.0, m, seps[10])
            n += 1
        dt = time.perf_counter() - t1
        per_point = dt / n * len(mus) / len(ls)      # 64 (mu) evaluations make 3 multipoles
        res["cpu_baseline"] = {"value": 1.0 / per_point, "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle (numpy, same rules) at 8 mu nodes of one separation "
                                         f"({dt:.1f} s), scaled to the 64-node projection"}
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nspec", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="batched: pick the workgroup layout and each length's radix order by timing them "
                         "on the batch first (coffe_amd/autotune.py); default: the stored wisdom file")
    ap.add_argument("--write-wisdom", default="", help="with --autotune: write the choice to this JSON file")
    ap.add_argument("--no-wisdom", action="store_true", help="batched: the engine's default layout and radix order")
    ap.add_argument("--graph", action="store_true",
                    help="batched: replay the step captured in a hipGraph (measured slower than the eager "
                         "call: 0.167 against 0.163 ms)")
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--workload", choices=["batched", "fullsky", "covariance", "divergent", "integrated"],
                    default="batched",
                    help="batched: the headline forward+adjoint batch; fullsky / covariance: the other "
                         "configurations, reported on their own lines")
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
              "each GPU would take a slice of the batch). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "fullsky":
        return run_fullsky(args)
    if args.workload == "covariance":
        return run_covariance(args)
    if args.workload == "divergent":
        return run_divergent(args)
    if args.workload == "integrated":
        return run_integrated(args)
    import torch
    from coffe_amd.engine import Batch, Engine
    import oracle as O

    specs, sep, A_np, bg, phys = build_inputs(args.nspec)
    nspec, nsep = len(specs), len(sep)
    n_in = [len(s[0]) for s in specs]
    n_fft = n_in
    eng = Engine(0)

    def make_batch():
        return Batch(eng, n_in, n_fft, [s[0][0] for s in specs], [s[0][-1] for s in specs], TABLES, sep, 1, len(LS))

    batch = make_batch()
    dev = "cuda"
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device=dev)
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device=dev)
    A = torch.tensor(A_np, device=dev)
    rng = np.random.default_rng(1)
    gout = torch.tensor(rng.standard_normal((nspec, 1, len(LS), nsep)), device=dev)
    itab = torch.empty((nspec, len(TABLES), nsep), dtype=torch.float64, device=dev)
    out = torch.empty((nspec, 1, len(LS), nsep), dtype=torch.float64, device=dev)
    gpk = torch.empty_like(p)
    tuned = None
    from coffe_amd import autotune as AT
    if args.autotune:
        # workgroup layout and per-length radix order measured on this batch
        batch.close()
        batch, rep = AT.autotune(make_batch, k, p, A, gout, n_fft)
        tuned = {"source": "measured in this run", "report": rep, "wisdom": AT.wisdom_of(rep)}
        if args.write_wisdom:
            with open(args.write_wisdom, "w") as f:
                json.dump(AT.wisdom_of(rep), f, indent=1)
    elif not args.no_wisdom and AT.load_wisdom() is not None:
        # the stored choice of an earlier autotune run on this batch shape
        batch.close()
        batch = AT.apply_wisdom(make_batch, AT.load_wisdom())
        tuned = {"source": "coffe_amd/wisdom_mi355x.json", "wisdom": AT.load_wisdom()}
    stream = torch.cuda.current_stream()
    sp = __import__("ctypes").c_void_p(stream.cuda_stream)
    kernels = ["k_resample", "k_prepare", "k_tables", "k_adj_c2r", "k_spline_adj"]
    launches = [lambda: batch.kernel_resample(k, p, sp), lambda: batch.kernel_fft_prepare(sp),
                lambda: batch.stage_tables(A, gout, itab, sp), lambda: batch.kernel_adj_c2r(sp),
                lambda: batch.kernel_spline_adj(k, gpk, sp)]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    kernel_ms = {kk: 0.0 for kk in kernels}

    def step():
        # one forward + adjoint pass over the batch: a single C call launches the five
        # kernels back to back on the stream (resample, prepare, tables, adj_c2r with
        # the projection to multipoles, spline_adj)
        batch.vjp(k, p, A, gout, itab=itab, out=out, gpk=gpk, stream=sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    run = step
    if args.graph:
        # the step captured once in a hipGraph (its five launches on the capture
        # stream) and replayed: no per-kernel launch work in the timed loop
        side = torch.cuda.Stream()
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            batch.vjp(k, p, A, gout, itab=itab, out=out, gpk=gpk,
                      stream=__import__("ctypes").c_void_p(side.cuda_stream))
        stream.wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            cs = torch.cuda.current_stream()
            batch.vjp(k, p, A, gout, itab=itab, out=out, gpk=gpk,
                      stream=__import__("ctypes").c_void_p(cs.cuda_stream))
        torch.cuda.synchronize()
        run = graph.replay
        for _ in range(max(args.warmup, 1)):
            run()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
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
    traffic = traffic_of("batched_256spec", dom)
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
                   "threads_per_wg": int(os.environ.get("COFFE_ENGINE_THREADS", 512)),
                   "launch": "hipGraph replay of the captured step" if args.graph else "eager (one C call)",
                   "autotune": tuned,
                   "other_configs": "bench.py --workload fullsky | covariance | divergent | integrated"},
        "roofline": {"bound": "mfma", "kernel": dom, "unit_kind": "FP64 VALU",
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
