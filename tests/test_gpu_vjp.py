"""GPU adjoint (VJP) of the batched multipoles against central finite differences of
the oracle: <gout, dM/dP v> from the engine's gpk must equal
<gout, (M(P + eps v) - M(P - eps v)) / (2 eps)> computed on the CPU."""
import numpy as np
import pytest

import oracle as O
from tests.synth import batch_spectra, flat_weights

pytestmark = pytest.mark.gpu


def _oracle_M(specs, tables, sep, A):
    tab = O.tables_at_sep_batch(specs, [len(s[0]) for s in specs], tables, sep)
    return np.einsum("zltr,str->szlr", A, tab)


@pytest.mark.parametrize("tables", [[(0, 0), (0, 2), (0, 4)], [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0),
                                                               (2, 2), (3, 1)]])
def test_vjp_matches_oracle_fd(engine, tables):
    import torch
    from coffe_amd.engine import Batch
    specs = [(s[0], s[1]) for s in batch_spectra(4, seed=7, lengths=(1024, 2048, 4096))]
    sep = np.linspace(5, 300, 48) / 0.67
    ntab = len(tables)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((2, 3, ntab, len(sep)))
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], tables, sep, 2, 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    gout_np = rng.standard_normal((len(specs), 2, 3, len(sep)))
    print("vjp ntab", ntab, flush=True)
    itab, out, gpk = b.vjp(k, p, torch.tensor(A, device="cuda"), torch.tensor(gout_np, device="cuda"))
    gpk = gpk.cpu().numpy()
    out = out.cpu().numpy()
    M0 = _oracle_M(specs, tables, sep, A)
    assert np.max(np.abs(out - M0)) <= 1e-10 * np.max(np.abs(M0))
    for trial in range(3):
        # smooth relative perturbations (Gaussian bumps in ln k): white noise at high k
        # puts grid-scale wiggles into the tables, where Akima's slope weights leave
        # their linear regime long before the FD step is small
        xi = []
        for s in specs:
            lk = np.log(s[0])
            c = rng.uniform(lk[0], lk[-1], 5)
            a = rng.standard_normal(5)
            xi.append(np.sum(a[:, None] * np.exp(-0.5 * ((lk[None, :] - c[:, None]) / 0.7) ** 2), axis=0))
        v = np.concatenate([s[1] * x for s, x in zip(specs, xi)])
        eps = 1e-6
        plus = [(s[0], s[1] * (1 + eps * x)) for s, x in zip(specs, xi)]
        minus = [(s[0], s[1] * (1 - eps * x)) for s, x in zip(specs, xi)]
        fd = np.sum(gout_np * (_oracle_M(plus, tables, sep, A) - _oracle_M(minus, tables, sep, A))) / (2 * eps)
        an = np.dot(gpk, v)
        assert abs(an - fd) <= 1e-6 * abs(fd), (trial, an, fd)


def test_vjp_homogeneity(engine):
    """The map P -> M is positively homogeneous of degree 1 (spline, FFTs and Akima
    all are), so <gpk, P> = <gout, M> exactly up to rounding."""
    import torch
    from coffe_amd.engine import Batch
    specs = [(s[0], s[1]) for s in batch_spectra(3, seed=9)]
    sep = np.linspace(10, 200, 64) / 0.67
    tables = [(0, 0), (0, 2), (0, 4)]
    A = flat_weights([0, 2, 4], 1.5, 0.87, 0.6, len(sep))
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], tables, sep, 1, 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    gout = torch.tensor(np.random.default_rng(2).standard_normal((3, 1, 3, len(sep))), device="cuda")
    _, out, gpk = b.vjp(k, p, torch.tensor(A, device="cuda"), gout)
    lhs = float(torch.dot(gpk, p))
    rhs = float(torch.sum(gout * out))
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)


def test_packed_workgroups_bit_identical(engine, monkeypatch):
    """Table units of one length share a workgroup (one group of threads and one LDS
    region each); with COFFE_ENGINE_PACK=0 every unit has its own workgroup. The
    arithmetic per unit is the same, so both give bit-identical tables, multipoles and
    gradients (lengths repeated so that packs of 2 and 4 units form)."""
    import torch
    from coffe_amd.engine import Batch
    specs = [(s[0], s[1]) for s in batch_spectra(5, seed=3, lengths=(1024, 2048))]
    tables = [(0, 0), (0, 2), (0, 4)]
    sep = np.linspace(5, 300, 48) / 0.67
    rng = np.random.default_rng(5)
    A = torch.tensor(rng.standard_normal((2, 3, 3, len(sep))), device="cuda")
    gout = torch.tensor(rng.standard_normal((len(specs), 2, 3, len(sep))), device="cuda")
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    args = ([len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
            [s[0][-1] for s in specs], tables, sep, 2, 3)
    res = []
    for pack in ("0", "1"):
        monkeypatch.setenv("COFFE_ENGINE_PACK", pack)
        b = Batch(engine, *args)
        print("pack", pack, flush=True)
        res.append([x.cpu().numpy() for x in b.vjp(k, p, A, gout)])
    monkeypatch.delenv("COFFE_ENGINE_PACK")
    for x0, x1 in zip(*res):
        assert np.array_equal(x0, x1)


@pytest.mark.parametrize("tables", [[(0, 0), (0, 2), (0, 4)], [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0),
                                                               (2, 2), (3, 1)]])
def test_vjp_exact_dot_product_white_noise(engine, tables):
    """The adjoint along arbitrary (white-noise) directions v of the samples, with no finite
    step: <gpk, v> must equal <gout, J v>, J v from the oracle's forward-mode derivative
    (the tables of v, then the dual-number Akima lookups), to 1e-9 of the sum of the two
    products' absolute terms (white noise puts intermediate FFT values far above the
    outputs, so rounding alone reaches ~1e-10 of it)."""
    import torch
    from coffe_amd.engine import Batch
    specs = [(s[0], s[1]) for s in batch_spectra(4, seed=7, lengths=(1024, 2048, 4096))]
    sep = np.linspace(5, 300, 48) / 0.67
    ntab = len(tables)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((2, 3, ntab, len(sep)))
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], tables, sep, 2, 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    gout = rng.standard_normal((len(specs), 2, 3, len(sep)))
    print(f"vjp dot product ntab={ntab}", flush=True)
    _, _, gpk = b.vjp(k, p, torch.tensor(A, device="cuda"), torch.tensor(gout, device="cuda"))
    gpk = gpk.cpu().numpy()
    for trial in range(3):
        dirs = [s[1] * rng.standard_normal(len(s[1])) for s in specs]
        _, tang = O.tables_at_sep_jvp(specs, [len(s[0]) for s in specs], tables, sep, dirs)
        Jv = np.einsum("zltr,str->szlr", A, tang)
        lhs = float(np.sum(gout * Jv))
        rhs = float(gpk @ np.concatenate(dirs))
        scale = float(np.sum(np.abs(gout * Jv))) + float(np.sum(np.abs(gpk * np.concatenate(dirs))))
        print(f"  trial {trial}: <gout, Jv> {lhs:.15e} <gpk, v> {rhs:.15e} scale {scale:.3e} "
              f"rel {abs(lhs - rhs) / scale:.2e}", flush=True)
        assert abs(lhs - rhs) <= 1e-9 * scale, (trial, lhs, rhs)
    b.close()
