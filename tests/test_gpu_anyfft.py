"""The general device FFT (coffe_fft_any: LDS lines, four-step, Bluestein) and the
FFTLog drop-in on it for the lengths the fused LDS kernels do not take:

- coffe_fft_any against numpy's FFT for lengths of each kind (direct 2^a 3^b 5^c <=
  8192, four-step products of two such lengths, Bluestein for primes and other
  factors), 1e-12 of the largest output;
- twofast_1bessel at N = 896 (N/2 = 7 * 64), 12002 (N/2 = 6001, a prime), 32768 and
  40000 against the oracle's restatement of twofast.c (its own mixed-radix and
  Bluestein FFT), 1e-10 of the largest output as the LDS path's tests;
- the Coffe object's integrals at integrals_fftlog_sampling = 40000 against the
  reference's fixtures (rtol 1e-4)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from tests.synth import spectrum

pytestmark = pytest.mark.gpu
DP = ctypes.POINTER(ctypes.c_double)
GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67


@pytest.mark.parametrize("L", [1, 2, 7, 60, 8192, 6000, 8193, 10000, 32768, 44100, 131071, 1000000])
def test_fft_any_vs_numpy(engine, L):
    from coffe_amd._lib import load_library
    Lb = load_library()
    rng = np.random.default_rng(L)
    x = rng.normal(size=L) + 1j * rng.normal(size=L)
    xin = np.ascontiguousarray(np.column_stack([x.real, x.imag]).reshape(-1))
    out = np.empty(2 * L)
    print(f"coffe_fft_any L={L}", flush=True)
    assert Lb.coffe_fft_any(xin.ctypes.data_as(DP), out.ctypes.data_as(DP), L) == 0
    got = out[0::2] + 1j * out[1::2]
    ref = np.fft.fft(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("N,l,nu", [(896, 0, 0.0), (12002, 2, 0.0), (32768, 0, 0.0), (40000, 1, 1.0)])
def test_twofast_general_lengths_vs_oracle(engine, N, l, nu):
    from coffe_amd.engine import twofast_1bessel
    k, pk = spectrum(4096)
    print(f"twofast_1bessel general path N={N} l={l} nu={nu}", flush=True)
    r, y = twofast_1bessel(k, pk, N, l, nu)
    ro, yo = O.twofast_1bessel(k, pk, N, l, nu)
    assert np.allclose(r, ro, rtol=1e-14, atol=0)
    assert np.max(np.abs(y - yo)) <= 1e-10 * np.max(np.abs(yo)), np.max(np.abs(y - yo)) / np.max(np.abs(yo))


def test_integrals_at_large_sampling(engine):
    from coffe_amd.coffe import Coffe
    COFFE_H0 = 1 / 2997.92458
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c = Coffe(integrals_fftlog_sampling=40000)
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    print("integrals at N = 40000", flush=True)
    for n, l in [(0, 0), (0, 2), (1, 1), (2, 2)]:
        idx = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)].index((n, l))
        xarr, yarr = np.loadtxt(os.path.join(GOLD, f"benchmark_integral{idx}.dat")).T
        for x, y in zip(xarr[::7], yarr[::7]):
            if 1.5 < x / (COFFE_H0 * H) < 20000:
                v = c.integral(r=x / (COFFE_H0 * H), l=l, n=n)
                if n > l:
                    v *= (COFFE_H0 * H) ** (n - l)
                assert np.isclose(y, v, rtol=1e-4), (n, l, x)
