"""Synthetic inputs of the benchmark configurations (seeded)."""
import numpy as np

H = 0.67


def pk_h(k, A=2e4, tilt=0.965, bao=105.0):
    """P(k) [Mpc/h]^3 at k [h/Mpc]: A k^n / (1 + (k/0.02)^2)^1.4 (1 + 0.05 sin(k s) e^{-(7k)^2})."""
    return A * k ** tilt / (1 + (k / 0.02) ** 2) ** 1.4 * (1 + 0.05 * np.sin(k * bao) * np.exp(-(7 * k) ** 2))


def spectrum(n, A=2e4, tilt=0.965, bao=105.0, kmin=1e-5, kmax=10.0):
    """n log-spaced samples in [kmin, kmax] h/Mpc, returned in Mpc units (k h, P / h^3)."""
    kh = np.logspace(np.log10(kmin), np.log10(kmax), n)
    return kh * H, pk_h(kh, A, tilt, bao) / H ** 3


def batch_spectra(nspec, seed=0, lengths=(1024, 2048, 4096, 8192, 16384)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nspec):
        A = rng.uniform(1.5e4, 2.5e4)
        tilt = rng.uniform(0.92, 1.0)
        bao = rng.uniform(95, 115)
        n = int(rng.choice(lengths))
        k, p = spectrum(n, A, tilt, bao)
        out.append((k, p, dict(A=A, tilt=tilt, bao=bao, n=n)))
    return out


def flat_weights(ls, b, f, D1, nsep, tables=((0, 0), (0, 2), (0, 4))):
    """mu-integrated weights of the flat-sky den+RSD integrand (functions.c:135-287,
    675-722 with the D1(z)^2 factor of 1457-1460) for tables I00, I02, I04:
    M_0 = D^2 (b^2 + 2bf/3 + f^2/5) I00, M_2 = -D^2 (4bf/3 + 4f^2/7) I02,
    M_4 = D^2 8f^2/35 I04."""
    A = np.zeros((1, len(ls), len(tables), nsep))
    coef = {0: (0, (b * b + 2 * b * f / 3 + f * f / 5)), 2: (1, -(4 * b * f / 3 + 4 * f * f / 7)),
            4: (2, 8 * f * f / 35)}
    for il, l in enumerate(ls):
        t, c = coef[l]
        A[0, il, t, :] = D1 * D1 * c
    return A
