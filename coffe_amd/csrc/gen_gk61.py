"""Generate the 61-point Gauss-Kronrod rule (QUADPACK qk61 layout) at high precision.

Kronrod extension of the 30-point Gauss-Legendre rule: the 31 added nodes are the
roots of the Stieltjes polynomial E_31, which is orthogonal to every polynomial of
degree <= 30 with respect to the weight P_30(x) on [-1, 1].  The weights follow from
exactness on even monomials up to degree 60 (odd ones vanish by symmetry), and the
rule is then checked for exactness up to degree 91.  Output order follows
QUADPACK's dqk61: xgk[0] is the largest node, odd indices are the Gauss nodes,
xgk[30] = 0.

Usage: python gen_gk61.py [output]  (default: gk61_nodes.h beside this file)
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 50
N = 30


def leg(n, x):
    """P_n(x) and P_n'(x) by the three-term recurrence."""
    p0, p1 = mp.mpf(1), x
    if n == 0:
        return p0, mp.mpf(0)
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1) if x * x != 1 else mp.mpf(0)
    return p1, dp


def gauss_legendre(n):
    xs, ws = [], []
    for i in range(1, n + 1):
        x = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))
        for _ in range(100):
            p, dp = leg(n, x)
            dx = p / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-mp.mp.dps + 5):
                break
        p, dp = leg(n, x)
        xs.append(x)
        ws.append(2 / ((1 - x * x) * dp * dp))
    return xs, ws  # descending


def main(out_path):
    gx, gw = gauss_legendre(N)
    gx_pos = [x for x in gx if x > 0]
    gw_pos = [w for x, w in zip(gx, gw) if x > 0]
    assert len(gx_pos) == 15
    qx, qw = gauss_legendre(80)

    def ip(f):
        return mp.fsum(w * f(x) for x, w in zip(qx, qw))

    basis = [N + 1 - 2 * k for k in range(0, 16)]  # 31, 29, ..., 1
    ms = list(range(1, N, 2))
    A = mp.matrix(len(ms), len(basis) - 1)
    b = mp.matrix(len(ms), 1)
    for i, m in enumerate(ms):
        for j, d in enumerate(basis[1:]):
            A[i, j] = ip(lambda x: leg(N, x)[0] * leg(m, x)[0] * leg(d, x)[0])
        b[i] = -ip(lambda x: leg(N, x)[0] * leg(m, x)[0] * leg(basis[0], x)[0])
    a = mp.lu_solve(A, b)

    def E(x):
        return leg(basis[0], x)[0] + mp.fsum(a[j] * leg(d, x)[0] for j, d in enumerate(basis[1:]))

    brackets = [(gx_pos[0], mp.mpf(1))] + [(gx_pos[i + 1], gx_pos[i]) for i in range(14)]
    kx = []
    for lo, hi in brackets:
        flo, fhi = E(lo), E(hi)
        assert flo * fhi < 0, "no sign change in bracket"
        for _ in range(200):
            mid = (lo + hi) / 2
            fm = E(mid)
            if fm * flo <= 0:
                hi, fhi = mid, fm
            else:
                lo, flo = mid, fm
        kx.append((lo + hi) / 2)
    xgk = []
    for i in range(15):
        xgk.append(kx[i])
        xgk.append(gx_pos[i])
    xgk.append(mp.mpf(0))
    W = mp.matrix(31, 31)
    rhs = mp.matrix(31, 1)
    for j in range(31):
        for i, x in enumerate(xgk):
            W[j, i] = (2 * x ** (2 * j)) if i < 30 else (mp.mpf(1) if j == 0 else mp.mpf(0))
        rhs[j] = mp.mpf(2) / (2 * j + 1)
    wgk = mp.lu_solve(W, rhs)
    # exactness check up to degree 90 (even) -- degree 91 odd is exact by symmetry
    for j in range(31, 46):
        s = mp.fsum(2 * w * x ** (2 * j) for x, w in zip(xgk[:30], wgk[:30]))
        assert abs(s - mp.mpf(2) / (2 * j + 1)) < mp.mpf(10) ** -40, j
    # the header the engine includes: the 31 Kronrod nodes xgk[0..30] with their weights,
    # and the 15 Gauss weights of the nodes xgk[1], xgk[3], ..., xgk[29]
    def fmt(v):
        """25 significant digits in d.ddd...eN form"""
        if v == 0:
            return "0.0"
        e10 = int(mp.floor(mp.log10(abs(v))))
        return "%se%d" % (mp.nstr(v / mp.mpf(10) ** e10, 25, strip_zeros=False, min_fixed=-1, max_fixed=2), e10)

    out = [
        "/* Kronrod nodes (descending, last = 0) and weights of the 61-point Gauss-Kronrod",
        "   rule; same values as QUADPACK dqk61 (generated at 50 digits, see gen_gk61.py). */",
        "static const double GK61K_X[31] = {%s};" % ", ".join(fmt(x) for x in xgk),
        "static const double GK61K_W[31] = {%s};" % ", ".join(fmt(w) for w in wgk),
        "/* 30-point Gauss weights of the nodes GK61K_X[1], GK61K_X[3], ..., GK61K_X[29] */",
        "static const double GK61G_W[15] = {%s};" % ", ".join(fmt(w) for w in gw_pos),
    ]
    with open(out_path, "w") as fh:
        fh.write("\n".join(out) + "\n")

if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "gk61_nodes.h"))
