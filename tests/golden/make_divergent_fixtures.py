"""Copies the reference's fixtures of the divergent table I_0^4 (tests/benchmarks/
benchmark_integral8.dat, 297 rows x = sep H0, y = I_0^4 sep^4 H0^4) and a decimated
copy of its renormalisation grid (benchmark_integral8_renormalization.dat, the
201 x 201 grid (x1, x2, value H0^4) reduced to every 10th node, 21 x 21 rows) into
tests/golden/.  Both come from COFFE's test suite (test_integrals.c:155-245); run
from the repository root with the reference checked out at /root/reference."""
import os

import numpy as np

SRC = "/root/reference/tests/benchmarks"
DST = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(os.path.join(SRC, "benchmark_integral8.dat")) as f:
        text = f.read()
    with open(os.path.join(DST, "benchmark_integral8.dat"), "w") as f:
        f.write(text)
    ren = np.loadtxt(os.path.join(SRC, "benchmark_integral8_renormalization.dat"))
    n = int(round(np.sqrt(len(ren))))
    assert n * n == len(ren)
    grid = ren.reshape(n, n, 3)
    keep = np.arange(0, n, 10)
    sub = grid[np.ix_(keep, keep)].reshape(-1, 3)
    np.savetxt(os.path.join(DST, "benchmark_integral8_renormalization_every10.dat"), sub, fmt="%.6e",
               header=f"every 10th node of the {n} x {n} grid of benchmark_integral8_renormalization.dat: x1 x2 value")


if __name__ == "__main__":
    main()
