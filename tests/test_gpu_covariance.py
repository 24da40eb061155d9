"""GPU parity of the 2D FFTLog covariance (cov_kernels.hip, covariance.c) against
the oracle's restatement of twobessel.c / covariance.c:
- the drop-in two_sph_bessel on diagonal inputs (the form covariance.c calls), full
  N x N output: 1e-11 of the largest |value|;
- the scaled D/G integrals at sep x sep (bicubic included), P and P^2, all pairs:
  1e-10 of the block's largest |value|;
- the Gaussian covariance assembly on the same integrals: 1e-12 relative;
- the whole GPU covariance at the reference default (N = 8000) against the reference's
  covariance fixtures, which were made with the STANDARD (direct quadrature) method
  (test_coffe.py:382-430): the two methods agree to 2e-3 here."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
LS = [0, 2, 4]


def _fixture_spectrum():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


# 1000 and 1200 = 2^4 * 3 * 5^2 take the radix-10 passes next to radix 2, 3, 4 and 5
@pytest.mark.parametrize("N", [64, 256, 1000, 1200, 2048])
@pytest.mark.parametrize("l12", [(0, 0), (0, 2), (2, 4), (4, 4)])
def test_two_sph_bessel_dropin(engine, N, l12):
    from coffe_amd.covariance import two_sph_bessel
    k = 1e-4 * (1e5) ** (np.arange(N) / N)
    d = k ** 3 * np.exp(-(k * 3.0) ** 2) * (1 + 0.1 * np.sin(40 * k)) / np.log(k[1] / k[0])
    fk = np.diag(d)
    print(f"two_sph_bessel N={N} l={l12}", flush=True)
    r1, r2, res = two_sph_bessel(k, k, fk, *l12)
    o1, o2, ores = O.two_sph_bessel(k, k, fk, *l12)
    assert np.array_equal(r1, o1) and np.array_equal(r2, o2)
    assert np.max(np.abs(res - ores)) <= 1e-11 * np.max(np.abs(ores))


def test_two_sph_bessel_rejects_non_diagonal(engine):
    from coffe_amd import EngineError
    from coffe_amd.covariance import two_sph_bessel
    k = 1e-3 * 1e3 ** (np.arange(64) / 64)
    fk = np.diag(k)
    fk[3, 5] = 1.0
    with pytest.raises(EngineError):
        two_sph_bessel(k, k, fk, 0, 0)


def _scaled(raw, l1, l2, power):
    return raw * (2 * l1 + 1) * (2 * l2 + 1) * (2.0 if power == 1 else 1.0) / np.pi / np.pi


@pytest.mark.parametrize("N", [1024])
def test_cov_integrals_vs_oracle(engine, N):
    import torch
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    plan = CovPlan(engine, N, k[0], k[-1], LS, sep)
    print(f"cov integrals N={N} rows={plan.nrows}", flush=True)
    got = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)]).cpu().numpy()
    for q, (i, j) in enumerate(plan.pairs):
        for power in (1, 2):
            ref = _scaled(O.covariance_fftlog_integral(k, pk, float(power), LS[i], LS[j], N, sep), LS[i], LS[j], power)
            g = got[0, power - 1, q]
            assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref)), (LS[i], LS[j], power)
    plan.close()


def test_cov_integrals_batch_vs_oracle(engine):
    import torch
    from coffe_amd.covariance import CovPlan
    from tests.synth import batch_spectra
    specs = batch_spectra(3, seed=4, lengths=(1024, 2048))
    kmin = max(s[0][0] for s in specs)
    kmax = min(s[0][-1] for s in specs)
    sep = np.array([30.0, 47.0, 90.0, 150.0, 260.0])
    plan = CovPlan(engine, 512, kmin, kmax, [0, 2], sep)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    got = plan.integrals(k, p, [len(s[0]) for s in specs]).cpu().numpy()
    for s, (ks, ps, _) in enumerate(specs):
        for q, (i, j) in enumerate(plan.pairs):
            for power in (1, 2):
                l1, l2 = [0, 2][i], [0, 2][j]
                ref = _scaled(O.covariance_fftlog_integral(ks, ps, float(power), l1, l2, 512, sep, kmin, kmax), l1, l2,
                              power)
                assert np.max(np.abs(got[s, power - 1, q] - ref)) <= 1e-10 * np.max(np.abs(ref)), (s, l1, l2, power)
    plan.close()


def _bin_inputs(bg, z=1.0, dz=0.1, fsky=0.2, pix=50 / H, nbar=1e-3 * H ** 3):
    chi = lambda zz: O.akima(bg["z"], bg["chi"], [zz])[0]  # noqa: E731
    vol = 4 * np.pi * fsky * (chi(z + dz) ** 3 - chi(z - dz) ** 3) / 3.
    f = O.akima(bg["z"], bg["f"], [z])[0]
    D = O.akima(bg["z"], bg["D1"], [z])[0]
    return dict(f=[f], D1=[D], nbar=[nbar], vol=[vol], pix=[pix])


def _oracle_blocks(integ, npair_pairs, nl, R):
    """GPU integrals [2][npair][R][R] -> the oracle's [nl*nl][R*R] blocks (i <= j)."""
    pk_int = np.zeros((nl * nl, R * R))
    pk2_int = np.zeros((nl * nl, R * R))
    for q, (i, j) in enumerate(npair_pairs):
        pk_int[i * nl + j] = integ[0, q].reshape(-1)
        pk2_int[i * nl + j] = integ[1, q].reshape(-1)
    return pk_int, pk2_int


def test_cov_assemble_vs_oracle(engine):
    import torch
    from coffe_amd import background as B
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    plan = CovPlan(engine, 512, k[0], k[-1], LS, sep)
    integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)])
    zi = _bin_inputs(bg)
    got = plan.assemble(integ, [1.3], **zi).cpu().numpy()[0]
    pk_int, pk2_int = _oracle_blocks(integ.cpu().numpy()[0], plan.pairs, 3, len(sep))
    ref = O.covariance_assemble(LS, sep, [1.3], zi["f"], zi["D1"], zi["nbar"], zi["vol"], zi["pix"], pk_int, pk2_int)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    plan.close()


@pytest.mark.parametrize("pops", [(1, 2, 1, 2), (1, 1, 2, 2), (2, 1, 2, 1), (1, 2, 2, 1), (3, 4, 1, 2)])
@pytest.mark.parametrize("window", [False, True])
def test_cov_assemble_populations_vs_oracle(engine, pops, window):
    """Multi-tracer covariance: populations, four distinct biases, density2 != density1,
    and the bin-averaged Poisson term, GPU vs oracle (covariance.c:790-1060)."""
    import torch
    from coffe_amd import background as B
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    plan = CovPlan(engine, 512, k[0], k[-1], LS, sep)
    print(f"cov populations {pops} window={window}", flush=True)
    integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)])
    zi = _bin_inputs(bg)
    bias = np.array([[1.3, 1.8, 0.9, 2.6]])
    nbar2 = [3.7e-4 * H ** 3]
    got = plan.assemble(integ, bias, pops=pops, nbar2=nbar2, window=window, **zi).cpu().numpy()[0]
    pk_int, pk2_int = _oracle_blocks(integ.cpu().numpy()[0], plan.pairs, 3, len(sep))
    ref = O.covariance_assemble(LS, sep, bias, zi["f"], zi["D1"], zi["nbar"], zi["vol"], zi["pix"], pk_int, pk2_int,
                                pops=pops, nbar2=nbar2, window=window)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    plan.close()


def test_cov_assemble_default_populations_read_bias1(engine):
    """With the reference's default populations (1, 1, 1, 1), galaxy biases 2..4 do not
    enter: the result equals the all-bias1 covariance (covariance.c:60-74)."""
    import torch
    from coffe_amd import background as B
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    plan = CovPlan(engine, 512, k[0], k[-1], LS, sep)
    integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)])
    zi = _bin_inputs(bg)
    many = plan.assemble(integ, np.array([[1.3, 2.1, 0.4, 3.3]]), **zi).cpu().numpy()
    one = plan.assemble(integ, [1.3], **zi).cpu().numpy()
    assert np.array_equal(many, one)
    plan.close()


def test_cov_vs_reference_fixtures(engine):
    """Default covariance settings of the reference test (z = 1, deltaz = 0.1, fsky = 0.2,
    pixel 50 Mpc/h, nbar = 1e-3 (h/Mpc)^3, b = 1, 2D FFTLog with N = 8000)."""
    import torch
    from coffe_amd import background as B
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    plan = CovPlan(engine, 8000, k[0], k[-1], LS, sep)
    print(f"cov fixtures N=8000 rows={plan.nrows}", flush=True)
    integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)])
    cov = plan.assemble(integ, [1.0], **_bin_inputs(bg)).cpu().numpy()[0, 0]
    plan.close()
    worst = 0.0
    for i, l1 in enumerate(LS):
        for j, l2 in enumerate(LS):
            x, y, z = np.loadtxt(os.path.join(GOLD, f"benchmark_multipoles_covariance_{l1}{l2}.dat")).T
            R = len(sep)
            got = cov[i, :, j, :].reshape(-1)   # (r1, r2) order as in the fixture
            assert np.allclose(np.repeat(sep, R) * H, x) and np.allclose(np.tile(sep, R) * H, y)
            rel = np.abs(got - z) / np.max(np.abs(z))
            worst = max(worst, float(np.max(rel)))
            print(f"  l1={l1} l2={l2} max rel (of block max) {np.max(rel):.2e}", flush=True)
    assert worst < 2e-3


def test_cov_integrals_n8000_vs_oracle(engine):
    """The covariance's default transform length (covariance_integration_bins = 8000, the
    bench's configuration) against the oracle's generic N x N transform and bicubic:
    the fixture spectrum, 16 separations 10-200 Mpc/h, pairs (0, 2) and (4, 4), P and P^2,
    1e-10 of each block's largest |value|."""
    import torch
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.linspace(10.0, 200.0, 16) / H
    plan = CovPlan(engine, 8000, k[0], k[-1], LS, sep)
    print(f"cov integrals N=8000 rows={plan.nrows}", flush=True)
    got = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)]).cpu().numpy()
    plan.close()
    for q, (i, j) in enumerate([(i, j) for i in range(3) for j in range(i, 3)]):
        if (LS[i], LS[j]) not in ((0, 2), (4, 4)):
            continue
        for power in (1, 2):
            print(f"  oracle N=8000 l=({LS[i]},{LS[j]}) power {power}", flush=True)
            ref = _scaled(O.covariance_fftlog_integral(k, pk, float(power), LS[i], LS[j], 8000, sep), LS[i], LS[j],
                          power)
            err = np.max(np.abs(got[0, power - 1, q] - ref)) / np.max(np.abs(ref))
            print(f"    max err (of block max) {err:.2e}", flush=True)
            assert err <= 1e-10, (LS[i], LS[j], power, err)
