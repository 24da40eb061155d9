"""The 2D FFTLog covariance at lengths outside the fused LDS row kernel (the general
device FFT: four-step for 10000 = 100 x 100 and 16384, Bluestein for 1792 = 2^8 7),
against the oracle's restatement of twobessel.c / covariance.c (its own mixed-radix and
Bluestein FFT):

- the drop-in two_sph_bessel on a diagonal input at N = 1792, full N x N output, 1e-11
  of the largest |value| (the LDS path's tolerance);
- the scaled D/G integrals at sep x sep, P and P^2, all pairs, at N = 1792 and 10000:
  1e-10 of each block's largest |value|;
- the adjoint at such lengths reports COFFE_EUNSUPPORTED."""
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


def _scaled(raw, l1, l2, power):
    return raw * (2 * l1 + 1) * (2 * l2 + 1) * (2.0 if power == 1 else 1.0) / np.pi / np.pi


@pytest.mark.parametrize("l12", [(0, 2), (4, 4)])
def test_two_sph_bessel_general_length(engine, l12):
    from coffe_amd.covariance import two_sph_bessel
    N = 1792
    k = 1e-4 * (1e5) ** (np.arange(N) / N)
    d = k ** 3 * np.exp(-(k * 3.0) ** 2) * (1 + 0.1 * np.sin(40 * k)) / np.log(k[1] / k[0])
    fk = np.diag(d)
    print(f"two_sph_bessel general N={N} l={l12}", flush=True)
    r1, r2, res = two_sph_bessel(k, k, fk, *l12)
    o1, o2, ores = O.two_sph_bessel(k, k, fk, *l12)
    assert np.array_equal(r1, o1) and np.array_equal(r2, o2)
    assert np.max(np.abs(res - ores)) <= 1e-11 * np.max(np.abs(ores)), np.max(np.abs(res - ores)) / np.max(np.abs(ores))


@pytest.mark.parametrize("N", [1792, 10000])
def test_cov_integrals_general_length_vs_oracle(engine, N):
    import torch
    from coffe_amd import EngineError
    from coffe_amd.covariance import CovPlan
    k, pk = _fixture_spectrum()
    sep = np.arange(50, 350, 50) / H
    plan = CovPlan(engine, N, k[0], k[-1], LS, sep)
    print(f"cov integrals general N={N} rows={plan.nrows}", flush=True)
    kt, pt = torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda")
    got = plan.integrals(kt, pt, [len(k)]).cpu().numpy()
    pairs = [(0, 2), (2, 2)] if N > 8192 else plan.pairs   # the oracle's 2D transform takes seconds at N = 10^4
    for (i, j) in pairs:
        q = plan.pairs.index((i, j))
        for power in (1, 2):
            ref = _scaled(O.covariance_fftlog_integral(k, pk, float(power), LS[i], LS[j], N, sep), LS[i], LS[j], power)
            g = got[0, power - 1, q]
            assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref)), (LS[i], LS[j], power)
    with pytest.raises(EngineError) as err:
        plan.integrals_vjp(kt, pt, [len(k)], torch.ones_like(torch.tensor(got, device="cuda")))
    assert err.value.code == 4
    plan.close()
