"""Autotuner of the batched engine's launch schedule, measured on the caller's own batch:
the table workgroups (units of one length packed into a workgroup or one per
workgroup, placed per XCD or in plain size order) and, for every FFT length of the
batch, the order of its radix passes.  Each candidate is timed as whole forward +
adjoint steps (coffe_batch_vjp) between HIP events on the launch stream; the fastest
is kept.  Changing the pass order changes results at rounding level only.

The choice can be stored as a wisdom file (JSON: the layout and the pass order per
transform length) and applied to a later batch without timing (`apply_wisdom`);
coffe_amd/wisdom_mi355x.json holds the one measured for the 256-spectrum batch of
lengths 1024-16384 on an MI355X.
"""
import ctypes
import itertools
import json
import os

import numpy as np

from ._lib import check


def radix_orders(M, limit=16):
    """Orders of radix passes for a transform of length M: the multisets of radices
    {8, 4, 2} with the fewest passes for the power of two (and the 5s and 3s of M),
    each in every distinct order, the engine's default order first."""
    a, m = 0, M
    while m % 2 == 0:
        m //= 2
        a += 1
    rest = []
    for r in (5, 3):
        while m % r == 0:
            m //= r
            rest.append(r)
    if m != 1:
        raise ValueError(f"length {M} has a prime factor above 5")
    p = -(-a // 3)
    sets = []
    for n8 in range(p, -1, -1):
        for n4 in range(p - n8, -1, -1):
            n2 = p - n8 - n4
            if 3 * n8 + 2 * n4 + n2 == a:
                sets.append([8] * n8 + [4] * n4 + [2] * n2 + rest)
    default = sets[0] if sets else rest
    out = [tuple(default)]
    for s in sets:
        for perm in sorted(set(itertools.permutations(s))):
            if perm not in out:
                out.append(perm)
            if len(out) >= limit:
                return out
    return out


def time_steps(batch, k, pk, A, gout, itab, out, gpk, reps=20, warmup=3):
    """Mean milliseconds of one coffe_batch_vjp step (HIP events on the current stream)."""
    import torch
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    for _ in range(warmup):
        batch.vjp(k, pk, A, gout, itab=itab, out=out, gpk=gpk, stream=sp)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(reps):
        batch.vjp(k, pk, A, gout, itab=itab, out=out, gpk=gpk, stream=sp)
    ev[1].record(stream)
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def set_radix(batch, M, order):
    arr = (ctypes.c_int * len(order))(*order)
    check(batch.L.coffe_batch_set_radix(batch.h, int(M), len(order), arr), "coffe_batch_set_radix")


WISDOM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wisdom_mi355x.json")


def load_wisdom(path=WISDOM):
    """The stored choice, or None when the file is absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def apply_wisdom(make_batch, wisdom):
    """Builds the batch with the stored layout and sets the stored pass order of every
    length the batch has (other lengths keep the default order)."""
    saved = {v: os.environ.get(v) for v in ("COFFE_ENGINE_PACK", "COFFE_ENGINE_XCD")}
    try:
        os.environ["COFFE_ENGINE_PACK"] = str(wisdom["layout"]["pack"])
        os.environ["COFFE_ENGINE_XCD"] = str(wisdom["layout"]["xcd"])
        batch = make_batch()
    finally:
        for v, x in saved.items():
            if x is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = x
    from coffe_amd import EngineError
    for M, order in wisdom["radix"].items():
        try:
            set_radix(batch, int(M), order)
        except EngineError as e:
            if e.code != 5:   # COFFE_EDOMAIN: no transform of that length in this batch
                raise
    return batch


def wisdom_of(report):
    return {"layout": {"pack": report["layout"]["pack"], "xcd": report["layout"]["xcd"]},
            "radix": {str(M): v["chosen"] for M, v in report["radix"].items()}}


def autotune(make_batch, k, pk, A, gout, n_fft, reps=20, radix_limit=16):
    """make_batch() builds the batch (it reads COFFE_ENGINE_PACK / COFFE_ENGINE_XCD at
    creation).  Returns (batch, report): the fastest workgroup layout, then per FFT
    length the fastest pass order, applied to the returned batch."""
    import torch
    nsep = gout.shape[-1]
    itab = torch.empty((gout.shape[0], A.shape[2], nsep), dtype=torch.float64, device=k.device)
    out = torch.empty_like(gout)
    gpk = torch.empty_like(pk)
    report = {"layouts": {}, "radix": {}}
    saved = {v: os.environ.get(v) for v in ("COFFE_ENGINE_PACK", "COFFE_ENGINE_XCD")}
    best = None
    try:
        for pack, xcd in (("1", "1"), ("1", "0"), ("0", "1")):
            os.environ["COFFE_ENGINE_PACK"], os.environ["COFFE_ENGINE_XCD"] = pack, xcd
            b = make_batch()
            ms = time_steps(b, k, pk, A, gout, itab, out, gpk, reps)
            report["layouts"][f"pack={pack},xcd={xcd}"] = ms
            if best is None or ms < best[0]:
                if best is not None:
                    best[1].close()
                best = (ms, b, pack, xcd)
            else:
                b.close()
    finally:
        for v, x in saved.items():
            if x is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = x
    ms_best, batch, pack, xcd = best
    report["layout"] = {"pack": pack, "xcd": xcd, "ms": ms_best}
    for M in sorted({int(n) // 2 for n in n_fft}, reverse=True):
        orders = radix_orders(M, radix_limit)
        times = []
        for order in orders:
            set_radix(batch, M, order)
            times.append(time_steps(batch, k, pk, A, gout, itab, out, gpk, reps))
        i = int(np.argmin(times))
        set_radix(batch, M, orders[i])
        report["radix"][M] = {"chosen": list(orders[i]), "ms": times[i], "default_ms": times[0],
                              "tried": len(orders)}
    report["ms"] = time_steps(batch, k, pk, A, gout, itab, out, gpk, reps)
    return batch, report
