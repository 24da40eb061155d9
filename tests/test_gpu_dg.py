"""The d2 and g1-g3 local terms on the GPU (k_mu_weights with the d2 / g sources: the
divergent table's weight and the renormalisation grid's bicubic at (chi1, chi2),
functions.c:574-669, 856-903, 1144-1428):

- GPU weights times the oracle's tables plus the GPU renormalisation part against the
  oracle's restatement (GL-256, the same divergent table and grid), every source alone
  and all together with density, RSD and Doppler-1 (their cross terms), full and flat
  sky, 1e-9 of each value plus 1e-12 of the monopole; with the full-sky RSD / Doppler
  terms the oracle's binary128 evaluation is the truth;
- the Coffe object with has_d2 / has_g1..3 against the reference's fixtures
  benchmark_{d2,g1,g2,g3}_multipoles (5e-4 of each value plus 1e-7 of the monopole:
  the reference's own np.allclose atol of 1e-8 exceeds every value);
- magnification and evolution biases through the Coffe object against the oracle."""
import os

import numpy as np
import pytest

import oracle as O
from coffe_amd import background as B

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


@pytest.fixture(scope="module")
def setup():
    from coffe_amd.engine import COFFE_SEP
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    zb = np.linspace(0, 10, 11)
    s1, s2, e1, e2 = 0.2 + 0.02 * zb, 0.3 - 0.01 * zb, 1.0 + 0.1 * zb, -0.5 + 0.05 * zb
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H, bias_z=zb, s1=s1, s2=s2, fevo1=e1, fevo2=e2)
    bgo = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"],
               a=bg["a"])
    tabs = {key: O.integral_table(key[0], key[1], k, pk, 4096, small_r=False) for key in O.TABLE_ORDER}
    chi_max = O.akima(bg["z"], bg["chi"], [2.3])[0]
    chi = np.arange(201) / 200 * chi_max
    i40, _, ren = O.divergent_tables(k, pk, COFFE_SEP, chi)
    return dict(bg=bg, bgo=bgo, tabs=tabs, zb=zb, mb=(s1, s2, e1, e2), div=(COFFE_SEP, i40), chi=chi, ren=ren)


@pytest.mark.parametrize("case", ["d2", "g1", "g2", "g3", "all", "all_flat"])
def test_dg_weights_vs_oracle(engine, setup, case):
    import torch
    from coffe_amd.engine import mu_weights
    S = setup
    srcs = {"d2", "g1", "g2", "g3"} if case.startswith("all") else {case}
    loc = dict(den=1, rsd=1, d1=1) if case.startswith("all") else dict(den=0, rsd=0, d1=0)
    flat = case == "all_flat"
    sep = np.array([20.0, 40.0, 80.0, 120.0, 200.0]) / H
    zm = [0.7, 1.5]
    ls = [0, 2, 4]
    b1, b2 = 1.2 + 0.3 * S["zb"], 1.5 + 0.1 * S["zb"]
    s1, s2, e1, e2 = S["mb"]
    Om, H0h = 0.30, H / 2997.92458
    dg = dict(sources=srcs, H0h=H0h, Omega_m=Om, mb_z=S["zb"], s1=s1, s2=s2, fevo1=e1, fevo2=e2, ren_chi=S["chi"],
              ren=S["ren"])
    tables = TABLES + [(4, 0)]
    print(f"d2/g weights {case}", flush=True)
    ren = torch.empty((len(zm), len(ls), len(sep)), dtype=torch.float64, device="cuda")
    A = mu_weights(engine, S["bgo"], S["zb"], b1, b2, zm, ls, tables, sep, flatsky=flat, nmu=256, dg=dg, ren_out=ren,
                   **loc).cpu().numpy()
    I = np.array([O.akima(*(S["tabs"][t] if t != (4, 0) else S["div"]), sep) for t in tables])
    got = np.einsum("zltr,tr->zlr", A, I) + ren.cpu().numpy()
    odg, keep = O.make_dg({k_: 1 for k_ in srcs}, H0h, Om, S["bg"]["a"], S["zb"], s1, s2, e1, e2, *S["div"], S["chi"],
                          S["ren"])
    if any(loc.values()) and not flat:
        # with the full-sky RSD and Doppler terms the truth is the binary128 evaluation of
        # the reference's formulas (an FP64 one carries ~1e-9 of (chi/s)^4 noise here)
        Ah, rh = O.mu_weights_hp(S["bgo"], S["zb"], b1, b2, zm, ls, sep, flatsky=flat, nmu=256, dg=odg, **loc)
        ref = np.einsum("zltr,tr->zlr", Ah, I) + rh
    else:
        ref = O.multipoles(S["tabs"], S["bgo"], S["zb"], b1, b2, zm, ls, sep, flatsky=flat, quad=256, dg=odg, **loc)
    for il, l in enumerate(ls):
        err = np.abs(got[:, il] - ref[:, il])
        bound = 1e-9 * np.abs(ref[:, il]) + 1e-12 * np.abs(ref[:, 0])
        print(f"  l={l}: max rel {np.max(err / np.abs(ref[:, il])):.2e}", flush=True)
        assert np.all(err <= bound), (case, l, np.max(err / bound))


@pytest.mark.parametrize("pre", ["d2", "g1", "g2", "g3"])
def test_dg_coffe_vs_reference_fixtures(engine, pre):
    from coffe_amd.coffe import Coffe
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c = Coffe(sep=np.array([10, 20, 40, 100, 150]) / H, l=[0, 2, 4])
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    c.reset_contributions()
    setattr(c, f"has_{pre}", True)
    print(f"Coffe {pre} multipoles", flush=True)
    res = c.compute_multipoles_bulk()
    mono = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_multipoles0.dat"))[:, 1]
    for l in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_multipoles{l}.dat")).T
        rows = [r for r in res if r.l == l]
        assert np.allclose([r.r * H for r in rows], x)
        got = np.array([r.value for r in rows])
        err = np.abs(got - y)
        print(f"  l={l}: max rel {np.max(err / np.abs(y)):.2e}", flush=True)
        assert np.all(err <= 5e-4 * np.abs(y) + 1e-7 * np.abs(mono)), (pre, l, err / np.abs(y))


def test_dg_coffe_magnification_and_evolution(engine, setup):
    """All local terms with non-zero magnification and evolution biases through the Coffe
    object: the background's G1 / G2 and the d2 / g factors read them; against the oracle
    on the same background (computed with the same biases) and the engine's tables."""
    from coffe_amd.coffe import Coffe
    S = setup
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    sep = np.array([30.0, 60.0, 120.0]) / H
    c = Coffe(sep=sep, l=[0, 2, 4], z_mean=[1.2], deltaz=[0.2], has_density=True, has_rsd=True, has_d1=True,
              has_d2=True, has_g1=True, has_g3=True, integrals_fftlog_sampling=4096)
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    s1, s2, e1, e2 = S["mb"]
    c.set_magnification_bias1(S["zb"], s1)
    c.set_magnification_bias2(S["zb"], s2)
    c.set_evolution_bias1(S["zb"], e1)
    c.set_evolution_bias2(S["zb"], e2)
    assert abs(c.magnification_bias1(1.0) - (0.2 + 0.02)) < 1e-12
    print("Coffe with s and f_evo", flush=True)
    got = np.array([r.value for r in c.compute_multipoles_bulk()]).reshape(3, len(sep))
    bgt = c._bg_tables()
    tabs = {t: c._integral_tables()[t] for t in TABLES}
    chi, ren = c._ren_grid(c.comoving_distance(1.4))
    sep_t, i40 = c._divergent_table()
    odg, keep = O.make_dg(dict(d2=1, g1=1, g3=1), H / 2997.92458, c.omega_m, bgt["a"], S["zb"], s1, s2, e1, e2, sep_t,
                          i40, chi, ren)
    one = np.ones(16)
    Ah, rh = O.mu_weights_hp(bgt, np.linspace(0, 15, 16), one, one, [1.2], [0, 2, 4], sep, den=1, rsd=1, d1=1,
                             flatsky=False, nmu=256, dg=odg)
    I = np.array([O.akima(*(tabs[t] if t != (4, 0) else (sep_t, i40)), sep) for t in TABLES + [(4, 0)]])
    ref = (np.einsum("zltr,tr->zlr", Ah, I) + rh)[0]
    for il in range(3):
        err = np.abs(got[il] - ref[il])
        assert np.all(err <= 1e-9 * np.abs(ref[il]) + 1e-12 * np.abs(ref[0])), (il, err / np.abs(ref[il]))


@pytest.mark.parametrize("flat", [False, True])
def test_dg_mutual_only_vs_oracle(engine, setup, flat):
    """COFFE_TERM_DG_MUTUAL (only_cross_correlations with d2 / g sources): the products of
    the chosen sources without their autos (functions.c:574-669 skip d2-d2, g1-g1, ...),
    against the oracle with those autos left out; full sky with the sources alone, flat
    sky next to density, RSD and Doppler-1 (whose own terms stay), 1e-9."""
    import torch
    from coffe_amd.engine import TERM_DG_MUTUAL, mu_weights, terms_of
    S = setup
    srcs = {"d2", "g1", "g3"}
    loc = dict(den=1, rsd=1, d1=1) if flat else dict(den=0, rsd=0, d1=0)
    sep = np.array([20.0, 80.0, 200.0]) / H
    zm, ls = [0.7, 1.5], [0, 2]
    b1, b2 = 1.2 + 0.3 * S["zb"], 1.5 + 0.1 * S["zb"]
    s1, s2, e1, e2 = S["mb"]
    Om, H0h = 0.30, H / 2997.92458
    dg = dict(sources=srcs, H0h=H0h, Omega_m=Om, mb_z=S["zb"], s1=s1, s2=s2, fevo1=e1, fevo2=e2, ren_chi=S["chi"],
              ren=S["ren"])
    tables = TABLES + [(4, 0)]
    terms = terms_of(loc["den"], loc["rsd"], loc["d1"], True) | TERM_DG_MUTUAL
    print(f"d2/g mutual only, flat={flat}", flush=True)
    ren = torch.empty((len(zm), len(ls), len(sep)), dtype=torch.float64, device="cuda")
    A = mu_weights(engine, S["bgo"], S["zb"], b1, b2, zm, ls, tables, sep, flatsky=flat, nmu=256, dg=dg, ren_out=ren,
                   terms=terms).cpu().numpy()
    I = np.array([O.akima(*(S["tabs"][t] if t != (4, 0) else S["div"]), sep) for t in tables])
    got = np.einsum("zltr,tr->zlr", A, I) + ren.cpu().numpy()
    odg, keep = O.make_dg({k_: 1 for k_ in srcs}, H0h, Om, S["bg"]["a"], S["zb"], s1, s2, e1, e2, *S["div"], S["chi"],
                          S["ren"], only_cross=1)
    ref = O.multipoles(S["tabs"], S["bgo"], S["zb"], b1, b2, zm, ls, sep, flatsky=flat, quad=256, dg=odg, **loc)
    for il in range(len(ls)):
        err = np.abs(got[:, il] - ref[:, il])
        assert np.all(err <= 1e-9 * np.abs(ref[:, il]) + 1e-12 * np.abs(ref[:, 0])), (flat, il, err / np.abs(ref[:, il]))


def test_dg_only_cross_coffe(engine):
    """has_only_cross_correlations with density, d2 and g1 through the Coffe object: the
    full result minus each contribution alone (their autos), to 1e-10 of the monopole."""
    from coffe_amd.coffe import Coffe
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T

    def run(**flags):
        c = Coffe(sep=np.array([20.0, 60.0, 150.0]) / H, l=[0, 2], has_density=False)
        c.set_power_spectrum_linear(k * H, pk / H ** 3)
        c.set_parameters(**flags)
        return np.array([r.value for r in c.compute_multipoles_bulk()])
    print("only_cross with d2 / g1 via Coffe", flush=True)
    both = dict(has_density=True, has_d2=True, has_g1=True)
    cross = run(**both, has_only_cross_correlations=True)
    want = run(**both) - run(has_density=True) - run(has_d2=True) - run(has_g1=True)
    mono = np.abs(run(**both)[:3]).max()
    assert np.allclose(cross, want, rtol=0, atol=1e-10 * mono), (cross, want)
