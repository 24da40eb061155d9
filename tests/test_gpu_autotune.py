"""The launch-schedule autotuner (coffe_amd/autotune.py) and the radix-order override
(coffe_batch_set_radix): any pass order gives the same tables, multipoles and
gradients to rounding, and invalid orders are refused."""
import numpy as np
import pytest

from tests.synth import batch_spectra, flat_weights

pytestmark = pytest.mark.gpu
TABLES = [(0, 0), (0, 2), (0, 4)]


def _inputs(engine, nspec=10):
    import torch
    from coffe_amd.engine import Batch
    specs = batch_spectra(nspec, seed=21, lengths=(1024, 2048, 8192))
    sep = np.linspace(5, 300, 64) / 0.67
    n_in = [len(s[0]) for s in specs]
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), **dev)
    p = torch.tensor(np.concatenate([s[1] for s in specs]), **dev)
    A = torch.tensor(flat_weights([0, 2, 4], 1.5, 0.87, 0.6, len(sep)), **dev)
    gout = torch.tensor(np.random.default_rng(4).standard_normal((nspec, 1, 3, len(sep))), **dev)

    def make():
        return Batch(engine, n_in, n_in, [s[0][0] for s in specs], [s[0][-1] for s in specs], TABLES, sep, 1, 3)
    return make, k, p, A, gout, n_in


def _close(a, b, tol):
    return np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))


# Tolerances of the three vjp outputs (tables, multipoles, pk gradient) relative to
# their largest element.  A different radix order changes every FFT's rounding; the
# pk gradient sums the adjoint transform against spectra spanning twelve decades, and
# over all orders of the 1024-point transform it moves by up to 1.0e-12 of its maximum
# (the tables and multipoles by at most 2.7e-15).
TOLS = (1e-12, 1e-12, 4e-12)


def test_set_radix_orders(engine):
    from coffe_amd import EngineError
    from coffe_amd.autotune import set_radix
    make, k, p, A, gout, n_in = _inputs(engine)
    b = make()
    print("set_radix: default vs reversed orders", flush=True)
    ref = [x.cpu().numpy() for x in b.vjp(k, p, A, gout)]
    set_radix(b, 4096, (8, 8, 8, 8))
    set_radix(b, 512, (8, 8, 8))
    set_radix(b, 1024, (2, 8, 8, 8))
    got = [x.cpu().numpy() for x in b.vjp(k, p, A, gout)]
    for x, y, tol in zip(got, ref, TOLS):
        assert _close(x, y, tol)
    with pytest.raises(EngineError):
        set_radix(b, 1024, (8, 8, 8, 4))        # product 2048
    with pytest.raises(EngineError):
        set_radix(b, 1024, (7, 8, 8, 2))
    with pytest.raises(EngineError):
        set_radix(b, 3000, (8, 5, 5, 5, 3))     # no transform of that length
    b.close()


def test_autotune_keeps_results(engine):
    from coffe_amd.autotune import autotune
    make, k, p, A, gout, n_in = _inputs(engine)
    b0 = make()
    ref = [x.cpu().numpy() for x in b0.vjp(k, p, A, gout)]
    b0.close()
    print("autotune", flush=True)
    b, rep = autotune(make, k, p, A, gout, n_in, reps=3)
    assert set(rep["radix"]) == {512, 1024, 4096}
    assert rep["ms"] > 0 and len(rep["layouts"]) == 3
    got = [x.cpu().numpy() for x in b.vjp(k, p, A, gout)]
    for x, y, tol in zip(got, ref, TOLS):
        assert _close(x, y, tol)
    b.close()
