"""Copies the reference's own test fixtures for this path into tests/golden/.

Data only (inputs and expected outputs of COFFE's tests, tests/benchmarks/ and the
input power spectrum PkL_CLASS.dat); the integral tables are decimated to the rows
the reference test checks (1.5 < r < 20000 Mpc, test_coffe.py:196-233), every 10th.
Run once in a container where the reference is mounted:
    python tests/golden/make_golden.py /root/reference
"""
import os
import shutil
import sys

import numpy as np

H = 0.67
COFFE_H0 = 1 / 2997.92458


def main(ref):
    out = os.path.dirname(os.path.abspath(__file__))
    bench = os.path.join(ref, "tests", "benchmarks")
    shutil.copy(os.path.join(ref, "PkL_CLASS.dat"), os.path.join(out, "PkL_CLASS.dat"))
    for i in range(8):
        d = np.loadtxt(os.path.join(bench, f"benchmark_integral{i}.dat"))
        r = d[:, 0] / (COFFE_H0 * H)
        m = (r > 1.5) & (r < 20000)
        np.savetxt(os.path.join(out, f"benchmark_integral{i}.dat"), d[m][::10], fmt="%.10e",
                   header=f"decimated copy of tests/benchmarks/benchmark_integral{i}.dat (x = r H0 h, value)")
    d = np.loadtxt(os.path.join(bench, "benchmark_background.dat"))
    np.savetxt(os.path.join(out, "benchmark_background.dat"), d[::50], fmt="%.10e",
               header="decimated copy of tests/benchmarks/benchmark_background.dat: z a H cH cH' D1 f chi")
    for pre in ("den", "rsd", "d1", "std", "std_d1", "d2", "g1", "g2", "g3", "len", "g4", "g5", "std_len", "all",
                "flatsky_lensing_lensing", "flatsky_density_lensing"):
        for l in (0, 2, 4):
            shutil.copy(os.path.join(bench, f"benchmark_{pre}_multipoles{l}.dat"), out)
    for pre in ("len", "g4", "g5", "std_len", "all", "flatsky_lensing_lensing"):
        for i in range(5):
            shutil.copy(os.path.join(bench, f"benchmark_{pre}_corrfunc{i}.dat"), out)
    for l1 in (0, 2, 4):
        for l2 in (0, 2, 4):
            shutil.copy(os.path.join(bench, f"benchmark_multipoles_covariance_{l1}{l2}.dat"), out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
