"""The batched forward + adjoint step captured in a hipGraph (torch.cuda.CUDAGraph) and
replayed: every replay equals an eager call on the same inputs (tables and multipoles
bit for bit; the gradient to rounding, since the adjoint Akima scatter adds into LDS
with atomics from several wavefronts, in the order they arrive), with eager work and
synchronisations between the replays, and the replays follow input changes made in
place (the graph holds pointers, not values)."""
import numpy as np
import pytest

from tests.synth import batch_spectra, flat_weights

pytestmark = pytest.mark.gpu

TABLES = [(0, 0), (0, 2), (0, 4)]


def test_vjp_graph_replays(engine):
    import torch
    from coffe_amd.engine import Batch
    specs = batch_spectra(24, seed=11)
    sep = np.linspace(5, 300, 128) / 0.67
    n_in = [len(s[0]) for s in specs]
    b = Batch(engine, n_in, n_in, [s[0][0] for s in specs], [s[0][-1] for s in specs], TABLES, sep, 1, 3)
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), **dev)
    p = torch.tensor(np.concatenate([s[1] for s in specs]), **dev)
    A = torch.tensor(flat_weights([0, 2, 4], 1.5, 0.87, 0.6, len(sep)), **dev)
    gout = torch.tensor(np.random.default_rng(2).standard_normal((len(specs), 1, 3, len(sep))), **dev)
    itab = torch.empty((len(specs), 3, len(sep)), **dev)
    out = torch.empty((len(specs), 1, 3, len(sep)), **dev)
    gpk = torch.empty_like(p)

    def step():
        s = torch.cuda.current_stream()
        import ctypes
        b.vjp(k, p, A, gout, itab=itab, out=out, gpk=gpk, stream=ctypes.c_void_p(s.cuda_stream))

    def eager():
        step()
        torch.cuda.synchronize()
        return itab.clone(), out.clone(), gpk.clone()

    print("graph: warm-up and capture of coffe_batch_vjp", flush=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()                      # warm-up on the capture stream (kernel attributes set)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    torch.cuda.synchronize()
    scales = [1.0, 1.01, 0.97, 1.0]
    p0 = p.clone()
    for it, sc in enumerate(scales):
        print(f"graph: replay {it} (P scaled by {sc})", flush=True)
        p.copy_(p0 * sc)
        ref = eager()                       # eager call on the same inputs
        itab.zero_(); out.zero_(); gpk.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(itab, ref[0]), it
        assert torch.equal(out, ref[1]), it
        assert torch.max(torch.abs(gpk - ref[2])).item() <= 1e-13 * torch.max(torch.abs(ref[2])).item(), it
        # eager work between replays
        _ = (p * 2.0).sum().item()
